// Host scheduler of the sdbf engine: implements the C ABI of include/sdbfx.h
// on top of the gfx950 kernels (compiled in this translation unit).
//
// Digesting: inputs (packed at 256-byte aligned offsets in device memory) are
// cut into stream-mode tiles or block-mode dd-blocks; the tile kernels run,
// the ordered insertion pass builds stream-mode filters, and everything lands
// in the set's device arrays (filters, per-filter element counts, Hamming
// weights) in input order.
//
// Comparing: every digest pair is oriented on the host with the reference's
// swap rule (sdbf/sdbf_core.cc:349-359), pairs are grouped by target digest,
// and the compare kernel + ordered-sum kernel produce one score per pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sdbfx.h"
#include "sdbfx_kernels.hip"

namespace {

// Largest input volume digested by one device pass (bounds scratch memory:
// about 1.4 bytes per byte of stream-mode input).  Inputs from host memory
// are staged into device memory pass by pass; device-resident inputs take
// larger passes, since each pass ends with the ordered stream insertion,
// whose length is that of its largest input's serial chain.
constexpr uint64_t kPassBytes = 4ull << 30;
constexpr uint64_t kPassBytesDevice = 16ull << 30;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

}  // namespace

struct sdbfx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;   // bytes of scratch (the DevBufs below) allowed, 0 = no limit
    uint64_t used = 0;     // bytes of scratch held
    std::string err;
    bool prof = false;
    int cmp_method = SDBFX_CMP_AUTO;
    // compare autotuning (sdbfx_set_autotune): per shape class, the launch
    // configuration found fastest by timing every candidate on the class's
    // first call
    bool autotune = false;
    std::map<uint64_t, sdbfx_tune_entry> tuned;
    struct Ev {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Ev> pending;
    std::map<std::string, std::pair<double, uint64_t>> times;
    std::vector<hipEvent_t> evpool;
    DevBuf stage, blocks, tiles, tstate, units, files, cand, sfilt, sham, counts, gsrc, gdst, gcnt,
        glast, groups, pairs, partial, imax, scores, gbest, rmap, rcode, text, tdesc, thdr, cofirst, cocount, feat,
        imatch, irefs;
    // two pinned host buffers for staging caller memory to the device
    void* pin[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    bool pin_busy[2] = {false, false};
    int pin_next = 0;
    // pinned ring for the small uploads (work lists, descriptors); its
    // space is reused after ring_ev, recorded after the latest of them
    uint8_t* ring = nullptr;
    uint64_t ring_head = 0;
    hipEvent_t ring_ev = nullptr;
    // pinned buffer for the small read-backs (counts, element counts, scores)
    uint8_t* rb = nullptr;
};

struct sdbfx_set {
    int device = 0;
    std::vector<std::string> names;
    std::vector<uint64_t> orig, first;
    std::vector<uint32_t> bfc, last, dd, maxe, last_ec;
    uint64_t nfilt = 0;
    DevBuf filt, ecf, hamf;
    // column order of the digests as comparison targets (col_order_kernel),
    // built on first use: per digest whether it is built, and the event
    // after the last build
    DevBuf cord, ckey;
    // per digest: 1 when folded to 64 bytes (sdbf::fast, sdbfx_set_fold)
    std::vector<uint8_t> fold;
    mutable std::vector<uint8_t> cbuilt;
    mutable hipEvent_t cev = nullptr;
    mutable std::mutex cmu;
};

// An index: bloom_filter(size, 5, 0, 0.01) (sdbf/bloom_filter.cc:52-71) in
// device memory -- `size` bytes of bits, bit mask 8 * size - 1, and its
// element count.  `first` (one word per bit, all ~0 between calls) is the
// per-bit scratch of block-mode insertion, allocated on first use.
struct sdbfx_index {
    int device = 0;
    uint64_t size = 0;
    uint32_t mask = 0;
    uint32_t* bits = nullptr;
    unsigned long long* elem = nullptr;
    uint32_t* first = nullptr;
};

// The no-allocation block digest (sdbfx_digest_blocks_noalloc) keeps its
// work list -- one BlockDesc per dd-block -- in this array instead of in
// memory the library allocates.  Each GPU holds its own copy for the life of
// the process; host code reaches it only through the device address that
// hipGetSymbolAddress gives for each device (scratch_ptr), never by this
// name, which in host code is a stand-in, not a device address.
#define SDBFX_NOALLOC_BLOCKS 131072   // dd-blocks per call: 2 GiB of input at 16 KiB blocks
static __device__ __align__(256) uint8_t scratch[SDBFX_NOALLOC_BLOCKS * sizeof(BlockDesc)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Resolves the device address of `scratch` on `device`, which must be the
// current device; once per device (sdbfx_ctx_create calls it).
int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return -1;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device])
        return 0;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return -1;
    g_scratch_ptr[device] = p;
    return 0;
}

void* scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace

namespace {

// every scratch buffer of a context
std::vector<DevBuf*> scratch_bufs(sdbfx_ctx* ctx)
{
    return {&ctx->stage, &ctx->blocks, &ctx->tiles, &ctx->tstate, &ctx->units, &ctx->files, &ctx->cand,
            &ctx->sfilt, &ctx->sham, &ctx->counts, &ctx->gsrc, &ctx->gdst, &ctx->gcnt, &ctx->glast, &ctx->groups,
            &ctx->pairs, &ctx->partial, &ctx->imax, &ctx->scores, &ctx->gbest, &ctx->rmap, &ctx->rcode,
            &ctx->text, &ctx->tdesc, &ctx->thdr, &ctx->cofirst, &ctx->cocount, &ctx->feat, &ctx->imatch,
            &ctx->irefs};
}

int fail(sdbfx_ctx* ctx, const std::string& msg)
{
    if (ctx)
        ctx->err = msg;
    return -1;
}

#define HIPC(ctx, call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(ctx, std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

// grows a scratch buffer of the context to at least `bytes` (contents not
// kept), within the context's budget: SDBFX_EBUDGET when the scratch held
// would exceed it (nothing is freed or allocated then)
int ensure(sdbfx_ctx* ctx, DevBuf& b, size_t bytes)
{
    if (b.cap >= bytes && b.p)
        return 0;
    size_t cap = std::max<size_t>(bytes, 256);
    cap = (cap + 4095) & ~(size_t)4095;
    if (ctx->budget && ctx->used - b.cap + cap > ctx->budget) {
        ctx->err = "scratch of " + std::to_string(ctx->used - b.cap + cap) + " bytes exceeds the context's budget of " +
                   std::to_string(ctx->budget);
        return SDBFX_EBUDGET;
    }
    if (b.p) {
        HIPC(ctx, hipStreamSynchronize(ctx->stream));
        HIPC(ctx, hipFree(b.p));
        ctx->used -= b.cap;
        b.p = nullptr;
        b.cap = 0;
    }
    HIPC(ctx, hipMalloc(&b.p, cap));
    b.cap = cap;
    ctx->used += cap;
    return 0;
}

// grows a device buffer keeping its first `keep` bytes
int grow_keep(sdbfx_ctx* ctx, DevBuf& b, size_t bytes, size_t keep)
{
    if (b.cap >= bytes && b.p)
        return 0;
    size_t cap = std::max<size_t>(bytes, b.cap + b.cap / 2);
    cap = (std::max<size_t>(cap, 4096) + 4095) & ~(size_t)4095;
    void* np = nullptr;
    HIPC(ctx, hipMalloc(&np, cap));
    if (b.p) {
        if (keep)
            HIPC(ctx, hipMemcpyAsync(np, b.p, keep, hipMemcpyDeviceToDevice, ctx->stream));
        HIPC(ctx, hipStreamSynchronize(ctx->stream));
        HIPC(ctx, hipFree(b.p));
    }
    b.p = np;
    b.cap = cap;
    return 0;
}

constexpr uint64_t kRingBytes = 16ull << 20;
// Copies below this size are made by a copy kernel even between pinned host
// memory and the device; small transfers are padded to it (the buffers on
// both sides hold at least this much) so that they are DMA copies.
constexpr uint64_t kDmaMin = 32ull << 10;

// A host-to-device copy of host data through the pinned ring, as a DMA copy:
// a copy kernel (what a copy from pageable memory, or a small one, becomes)
// waits for compute units behind a comparison running on another stream, and
// holds up this stream and the calling thread with it.  dst must hold
// max(bytes, kDmaMin).  Copies larger than the ring go directly.
int upload_small(sdbfx_ctx* ctx, void* dst, const void* src, uint64_t bytes)
{
    const uint64_t xfer = std::max(bytes, kDmaMin);
    const uint64_t need = (xfer + 255) & ~255ull;
    if (need > kRingBytes) {
        HIPC(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return 0;
    }
    if (!ctx->ring) {
        HIPC(ctx, hipHostMalloc((void**)&ctx->ring, kRingBytes, hipHostMallocDefault));
        HIPC(ctx, hipEventCreate(&ctx->ring_ev));
        ctx->ring_head = 0;
    } else if (ctx->ring_head + need > kRingBytes) {
        // every earlier copy out of the ring is done once the latest is
        HIPC(ctx, hipEventSynchronize(ctx->ring_ev));
        ctx->ring_head = 0;
    }
    uint8_t* p = ctx->ring + ctx->ring_head;
    ctx->ring_head += need;
    memcpy(p, src, bytes);
    HIPC(ctx, hipMemcpyAsync(dst, p, xfer, hipMemcpyHostToDevice, ctx->stream));
    HIPC(ctx, hipEventRecord(ctx->ring_ev, ctx->stream));
    return 0;
}

constexpr uint64_t kRbBytes = 4ull << 20;

// A device-to-host read-back that waits for the stream, through the pinned
// read-back buffer so that it is a DMA copy (into pageable memory, or when
// small, it would be a copy kernel, held up behind a comparison on another
// stream); `avail` bytes are readable at src, and up to kDmaMin of them are
// read.  Larger read-backs go directly.
int download(sdbfx_ctx* ctx, void* dst, const void* src, uint64_t bytes, uint64_t avail)
{
    if (bytes > kRbBytes) {
        HIPC(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPC(ctx, hipStreamSynchronize(ctx->stream));
        return 0;
    }
    if (!ctx->rb)
        HIPC(ctx, hipHostMalloc((void**)&ctx->rb, kRbBytes, hipHostMallocDefault));
    if (bytes)
        HIPC(ctx, hipMemcpyAsync(ctx->rb, src, std::max(bytes, std::min(avail, kDmaMin)), hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(dst, ctx->rb, bytes);
    return 0;
}

template <typename T>
int upload(sdbfx_ctx* ctx, DevBuf& b, const std::vector<T>& v)
{
    if (int rc = ensure(ctx, b, std::max<uint64_t>(v.size() * sizeof(T) + 16, kDmaMin)))
        return rc;
    if (!v.empty())
        return upload_small(ctx, b.p, v.data(), v.size() * sizeof(T));
    return 0;
}

hipEvent_t get_event(sdbfx_ctx* ctx)
{
    if (!ctx->evpool.empty()) {
        hipEvent_t e = ctx->evpool.back();
        ctx->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

int collect_times(sdbfx_ctx* ctx)
{
    if (ctx->pending.empty())
        return 0;
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& ev : ctx->pending) {
        float ms = 0.f;
        HIPC(ctx, hipEventElapsedTime(&ms, ev.a, ev.b));
        auto& t = ctx->times[ev.name];
        t.first += ms;
        t.second += 1;
        ctx->evpool.push_back(ev.a);
        ctx->evpool.push_back(ev.b);
    }
    ctx->pending.clear();
    return 0;
}

// Launches a kernel on the context stream, bracketed by events when profiling.
#define LAUNCH(ctx, kname, grid, block, ...)                                               \
    do {                                                                                   \
        sdbfx_ctx::Ev ev_{#kname, nullptr, nullptr};                                       \
        if ((ctx)->prof) {                                                                 \
            ev_.a = get_event(ctx);                                                        \
            ev_.b = get_event(ctx);                                                        \
            HIPC(ctx, hipEventRecord(ev_.a, (ctx)->stream));                               \
        }                                                                                  \
        hipLaunchKernelGGL(kname, dim3(grid), dim3(block), 0, (ctx)->stream, __VA_ARGS__); \
        HIPC(ctx, hipGetLastError());                                                      \
        if ((ctx)->prof) {                                                                 \
            HIPC(ctx, hipEventRecord(ev_.b, (ctx)->stream));                               \
            (ctx)->pending.push_back(ev_);                                                 \
        }                                                                                  \
    } while (0)

int set_reserve(sdbfx_ctx* ctx, sdbfx_set* s, uint64_t total)
{
    if (int rc = grow_keep(ctx, s->filt, total * SDBFX_BF + 256, s->nfilt * SDBFX_BF))
        return rc;
    if (int rc = grow_keep(ctx, s->ecf, total * 2 + 256, s->nfilt * 2))
        return rc;
    if (int rc = grow_keep(ctx, s->hamf, total * 2 + 256, s->nfilt * 2))
        return rc;
    if (int rc = grow_keep(ctx, s->cord, total * 4 + 256, s->nfilt * 4))
        return rc;
    if (int rc = grow_keep(ctx, s->ckey, total * 4 + 256, s->nfilt * 4))
        return rc;
    return 0;
}

uint64_t block_count(uint64_t len, uint32_t dd)
{
    return len / dd + ((len % dd) >= SDBFX_MIN_FILE ? 1 : 0);
}

// The index_info of a digest call (sdbf/index_info.h): `ix` is info->index
// (stream-mode gate, block-mode insertion), `search` the indexes of
// info->setlist (block-mode lookups, which take the place of the insertion);
// matches receives nsearch counts per block of the call's block-mode
// inputs, `mpos` blocks written so far.
struct IndexCall {
    sdbfx_index* ix = nullptr;
    std::vector<IndexRef> search;
    uint32_t* matches = nullptr;
    uint64_t mpos = 0;
};

// Block-mode index work of one pass, after block_digest_kernel left the
// counted features in ctx->feat.
int block_index_pass(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t nblk, IndexCall* ic)
{
    const uint32_t nix = (uint32_t)ic->search.size();
    const BlockDesc* bl = (const BlockDesc*)ctx->blocks.p;
    const uint16_t* ecf = (const uint16_t*)set->ecf.p;
    const uint32_t* feat = (const uint32_t*)ctx->feat.p;
    if (nix) {
        int rc;
        if ((rc = ensure(ctx, ctx->imatch, (uint64_t)nblk * nix * 4 + 256)) || (rc = upload(ctx, ctx->irefs, ic->search)))
            return rc;
        HIPC(ctx, hipMemsetAsync(ctx->imatch.p, 0, (uint64_t)nblk * nix * 4, ctx->stream));
        const uint64_t nt = (uint64_t)nblk * (SDBFX_MAX_ELEM_DD / 4);
        LAUNCH(ctx, index_search_kernel, (uint32_t)((nt + 255) / 256), 256, bl, nblk, ecf, feat,
               (const IndexRef*)ctx->irefs.p, nix, (uint32_t*)ctx->imatch.p);
        if (int rc2 = download(ctx, ic->matches + ic->mpos * nix, ctx->imatch.p, (uint64_t)nblk * nix * 4,
                               ctx->imatch.cap))
            return rc2;
        ic->mpos += nblk;
        return 0;
    }
    sdbfx_index* ix = ic->ix;
    const uint64_t nbits = 8ull * ix->size;
    if (!ix->first) {
        HIPC(ctx, hipMalloc(&ix->first, nbits * 4));
        LAUNCH(ctx, fill_u32_kernel, (uint32_t)std::min<uint64_t>(2048, (nbits / 64 + 255) / 256), 256, ix->first,
               nbits, 0xFFFFFFFFu);
    }
    const uint64_t nt = (uint64_t)nblk * SDBFX_MAX_ELEM_DD;
    const uint32_t grid = (uint32_t)((nt + 255) / 256);
    LAUNCH(ctx, index_first_kernel, grid, 256, bl, nblk, ecf, feat, (const uint32_t*)ix->bits, ix->mask, ix->first);
    LAUNCH(ctx, index_commit_kernel, grid, 256, bl, nblk, ecf, feat, ix->bits, ix->mask,
           (const uint32_t*)ix->first, ix->elem);
    LAUNCH(ctx, index_reset_kernel, grid, 256, bl, nblk, ecf, feat, ix->mask, ix->first);
    return 0;
}

// One device pass over inputs [i0, i1) that are resident at d_base + off[i].
int digest_pass(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t i0, uint32_t i1, const uint8_t* d_base,
                const uint64_t* off, const uint64_t* lens, const char* const* names, const uint32_t* dd,
                const int32_t* status, IndexCall* ic)
{
    // ---- stream-mode work list
    std::vector<TileDesc> tiles;
    std::vector<UnitDesc> units;
    std::vector<StreamFile> sfiles;
    std::vector<uint32_t> sidx;   // input index of each stream file
    uint64_t sfilt_total = 0;
    for (uint32_t i = i0; i < i1; i++) {
        if (status[i] != 0 || (dd && dd[i]))
            continue;
        StreamFile sf{};
        sf.tile_begin = (uint32_t)tiles.size();
        const uint64_t len = lens[i];
        for (uint64_t u0 = 0; u0 < len; u0 += SDBFX_CHUNK) {
            const uint32_t n = (uint32_t)std::min<uint64_t>(SDBFX_CHUNK, len - u0);
            if (n <= 64)
                continue;
            const uint32_t tb = (uint32_t)tiles.size();
            for (uint32_t t0 = 0; t0 < n - 64; t0 += SDBFX_TILE) {
                TileDesc td{};
                td.unit_off = off[i] + u0;
                td.n = n;
                td.t0 = t0;
                td.T = std::min<uint32_t>(SDBFX_TILE, n - 64 - t0);
                td.first = t0 == 0;
                tiles.push_back(td);
            }
            if (tiles.size() - tb > 1)
                units.push_back(UnitDesc{tb, (uint32_t)tiles.size()});
        }
        sf.tile_end = (uint32_t)tiles.size();
        sf.filt_base = sfilt_total;
        sf.filt_cap = (uint32_t)((len >> 11) + 2);
        sfilt_total += sf.filt_cap;
        sfiles.push_back(sf);
        sidx.push_back(i);
    }

    std::vector<uint32_t> counts(3 * sfiles.size());
    if (!sfiles.empty()) {
        int rc;
        if ((rc = upload(ctx, ctx->tiles, tiles)) || (rc = upload(ctx, ctx->units, units)) ||
            (rc = upload(ctx, ctx->files, sfiles)))
            return rc;
        if ((rc = ensure(ctx, ctx->cand, tiles.size() * 5ull * SDBFX_CAND_CAP * 4 + 256)) ||
            (rc = ensure(ctx, ctx->tstate, tiles.size() * sizeof(TileState) + 256)) ||
            (rc = ensure(ctx, ctx->sfilt, sfilt_total * SDBFX_BF + 256)) ||
            (rc = ensure(ctx, ctx->sham, sfilt_total * 2 + 256)) ||
            (rc = ensure(ctx, ctx->counts, counts.size() * 4 + 256)))
            return rc;
        if (!tiles.empty())
            LAUNCH(ctx, stream_tile_kernel, (uint32_t)tiles.size(), SDBFX_NT, d_base,
                   (const TileDesc*)ctx->tiles.p, (uint32_t*)ctx->cand.p, (TileState*)ctx->tstate.p);
        if (!units.empty()) {
            LAUNCH(ctx, stream_fixup_plan_kernel, (uint32_t)((tiles.size() + 255) / 256), 256,
                   (const TileDesc*)ctx->tiles.p, (uint32_t)tiles.size(), (TileState*)ctx->tstate.p);
            LAUNCH(ctx, stream_fixup_par_kernel, (uint32_t)tiles.size(), SDBFX_NT, d_base,
                   (const TileDesc*)ctx->tiles.p, (uint32_t*)ctx->cand.p, (TileState*)ctx->tstate.p);
            LAUNCH(ctx, stream_fixup_kernel, (uint32_t)units.size(), SDBFX_NT, d_base,
                   (const TileDesc*)ctx->tiles.p, (const UnitDesc*)ctx->units.p, (uint32_t*)ctx->cand.p,
                   (TileState*)ctx->tstate.p);
        }
        if (ic && ic->ix)
            LAUNCH(ctx, stream_insert_indexed_kernel, 1, 64, (const StreamFile*)ctx->files.p, (uint32_t)sfiles.size(),
                   (const TileDesc*)ctx->tiles.p, (const TileState*)ctx->tstate.p, (const uint32_t*)ctx->cand.p,
                   (uint8_t*)ctx->sfilt.p, (uint16_t*)ctx->sham.p, (uint32_t*)ctx->counts.p, ic->ix->bits,
                   ic->ix->mask, ic->ix->elem);
        else
            LAUNCH(ctx, stream_insert_kernel, (uint32_t)sfiles.size(), 64, (const StreamFile*)ctx->files.p,
                   (const TileDesc*)ctx->tiles.p, (const TileState*)ctx->tstate.p,
                   (const uint32_t*)ctx->cand.p, (uint8_t*)ctx->sfilt.p, (uint16_t*)ctx->sham.p,
                   (uint32_t*)ctx->counts.p);
        if ((rc = download(ctx, counts.data(), ctx->counts.p, counts.size() * 4, ctx->counts.cap)))
            return rc;
        for (size_t f = 0; f < sfiles.size(); f++)
            if (counts[3 * f + 2])
                return fail(ctx, "stream digest exceeded its filter capacity");
    }

    // ---- final filter ranges in input order
    uint64_t total = 0;
    std::vector<uint64_t> first(i1 - i0, 0), cnt(i1 - i0, 0);
    {
        size_t f = 0;
        for (uint32_t i = i0; i < i1; i++) {
            if (status[i] != 0)
                continue;
            uint64_t c;
            if (dd && dd[i]) {
                c = block_count(lens[i], dd[i]);
            } else {
                c = counts[3 * f];
                f++;
            }
            first[i - i0] = set->nfilt + total;
            cnt[i - i0] = c;
            total += c;
        }
    }
    if (int rc = set_reserve(ctx, set, set->nfilt + total))
        return rc;

    // ---- block mode straight into the set
    std::vector<BlockDesc> blocks;
    for (uint32_t i = i0; i < i1; i++) {
        if (status[i] != 0 || !dd || !dd[i])
            continue;
        const uint64_t len = lens[i];
        const uint32_t d = dd[i];
        const uint64_t qt = len / d, rem = len % d;
        for (uint64_t b = 0; b < qt; b++)
            blocks.push_back(BlockDesc{off[i] + b * d, d, 0, first[i - i0] + b});
        if (rem >= SDBFX_MIN_FILE)
            blocks.push_back(BlockDesc{off[i] + qt * d, (uint32_t)rem, 1, first[i - i0] + qt});
    }
    if (!blocks.empty()) {
        if (int rc = upload(ctx, ctx->blocks, blocks))
            return rc;
        const bool idx = ic && (ic->ix || !ic->search.empty());
        if (idx)
            if (int rc = ensure(ctx, ctx->feat, blocks.size() * (SDBFX_MAX_ELEM_DD * 5 * 4ull) + 256))
                return rc;
        LAUNCH(ctx, block_digest_kernel, (uint32_t)blocks.size(), SDBFX_NT, d_base,
               (const BlockDesc*)ctx->blocks.p, (uint8_t*)set->filt.p, (uint16_t*)set->ecf.p,
               (uint16_t*)set->hamf.p, idx ? (uint32_t*)ctx->feat.p : (uint32_t*)nullptr);
        if (idx)
            if (int rc = block_index_pass(ctx, set, (uint32_t)blocks.size(), ic))
                return rc;
    }

    // ---- stream filters into the set
    if (!sfiles.empty()) {
        std::vector<uint64_t> gsrc(sfiles.size()), gdst(sfiles.size());
        std::vector<uint32_t> gcnt(sfiles.size()), glast(sfiles.size());
        for (size_t f = 0; f < sfiles.size(); f++) {
            gsrc[f] = sfiles[f].filt_base;
            gdst[f] = first[sidx[f] - i0];
            gcnt[f] = counts[3 * f];
            glast[f] = counts[3 * f + 1];
        }
        int rc;
        if ((rc = upload(ctx, ctx->gsrc, gsrc)) || (rc = upload(ctx, ctx->gdst, gdst)) ||
            (rc = upload(ctx, ctx->gcnt, gcnt)) || (rc = upload(ctx, ctx->glast, glast)))
            return rc;
        LAUNCH(ctx, gather_filters_kernel, (uint32_t)sfiles.size(), 64, (const uint8_t*)ctx->sfilt.p,
               (const uint16_t*)ctx->sham.p, (const uint64_t*)ctx->gsrc.p, (const uint64_t*)ctx->gdst.p,
               (const uint32_t*)ctx->gcnt.p, (const uint32_t*)ctx->glast.p, (uint8_t*)set->filt.p,
               (uint16_t*)set->hamf.p, (uint16_t*)set->ecf.p);
    }

    // ---- metadata (block digests need their last element count on the host)
    std::vector<uint16_t> ec_all(total);
    if (total) {
        if (int rc = download(ctx, ec_all.data(), (uint16_t*)set->ecf.p + set->nfilt, total * 2,
                              set->ecf.cap - set->nfilt * 2))
            return rc;
    }
    size_t f = 0;
    for (uint32_t i = i0; i < i1; i++) {
        if (status[i] != 0)
            continue;
        const bool blk = dd && dd[i];
        set->names.push_back(names && names[i] ? names[i] : "");
        set->fold.push_back(0);
        set->orig.push_back(lens[i]);
        set->first.push_back(first[i - i0]);
        set->bfc.push_back((uint32_t)cnt[i - i0]);
        set->dd.push_back(blk ? dd[i] : 0);
        set->maxe.push_back(blk ? SDBFX_MAX_ELEM_DD : SDBFX_MAX_ELEM);
        if (blk) {
            set->last.push_back(0);
            set->last_ec.push_back(ec_all[first[i - i0] + cnt[i - i0] - 1 - set->nfilt]);
        } else {
            set->last.push_back(counts[3 * f + 1]);
            set->last_ec.push_back(counts[3 * f + 1]);
            f++;
        }
    }
    set->nfilt += total;
    return 0;
}

constexpr uint64_t kPinBytes = 8ull << 20;   // per pinned staging buffer
constexpr uint64_t kPinMin = 64ull << 10;    // smaller copies go directly

// Copies caller (pageable) memory to the device through the two pinned
// buffers: one is filled on the host while the other's copy is in flight.
int upload_host(sdbfx_ctx* ctx, uint8_t* dst, const uint8_t* src, uint64_t len)
{
    if (len < kPinMin) {
        HIPC(ctx, hipMemcpyAsync(dst, src, len, hipMemcpyHostToDevice, ctx->stream));
        return 0;
    }
    if (!ctx->pin[0])
        for (int k = 0; k < 2; k++) {
            HIPC(ctx, hipHostMalloc(&ctx->pin[k], kPinBytes, hipHostMallocDefault));
            HIPC(ctx, hipEventCreate(&ctx->pin_ev[k]));
        }
    for (uint64_t o = 0; o < len; o += kPinBytes) {
        const int k = ctx->pin_next;
        ctx->pin_next ^= 1;
        if (ctx->pin_busy[k])
            HIPC(ctx, hipEventSynchronize(ctx->pin_ev[k]));
        const uint64_t c = std::min(kPinBytes, len - o);
        memcpy(ctx->pin[k], src + o, c);
        HIPC(ctx, hipMemcpyAsync(dst + o, ctx->pin[k], c, hipMemcpyHostToDevice, ctx->stream));
        HIPC(ctx, hipEventRecord(ctx->pin_ev[k], ctx->stream));
        ctx->pin_busy[k] = true;
    }
    return 0;
}

// The padded device image of a pass's host inputs, bytes [0, end) of dst:
// input i at hoff[i] (i in [i0, i1)), zeros elsewhere.  It is assembled on
// the host in windows of the two pinned buffers, each window one copy, so
// that every transfer is a DMA copy from pinned memory: nothing on the
// stream needs compute units (a copy kernel or a fill would wait for them
// behind a comparison running on another stream, and the staging with it).
int upload_pass(sdbfx_ctx* ctx, uint8_t* dst, uint32_t i0, uint32_t i1, const std::vector<uint64_t>& hoff,
                const uint8_t* const* host_bufs, const uint64_t* lens, uint64_t end)
{
    if (!ctx->pin[0])
        for (int k = 0; k < 2; k++) {
            HIPC(ctx, hipHostMalloc(&ctx->pin[k], kPinBytes, hipHostMallocDefault));
            HIPC(ctx, hipEventCreate(&ctx->pin_ev[k]));
        }
    uint32_t j = i0;   // the first input that ends after the window's start
    for (uint64_t w0 = 0; w0 < end; w0 += kPinBytes) {
        const uint64_t w1 = std::min(end, w0 + kPinBytes);
        const int k = ctx->pin_next;
        ctx->pin_next ^= 1;
        if (ctx->pin_busy[k])
            HIPC(ctx, hipEventSynchronize(ctx->pin_ev[k]));
        uint8_t* pb = (uint8_t*)ctx->pin[k];
        uint64_t c = w0;
        while (c < w1) {
            while (j < i1 && hoff[j] + lens[j] <= c)
                j++;
            if (j == i1 || hoff[j] >= w1) {
                memset(pb + (c - w0), 0, w1 - c);
                break;
            }
            if (c < hoff[j]) {
                memset(pb + (c - w0), 0, hoff[j] - c);
                c = hoff[j];
            }
            const uint64_t e = std::min(w1, hoff[j] + lens[j]);
            memcpy(pb + (c - w0), host_bufs[j] + (c - hoff[j]), e - c);
            c = e;
        }
        HIPC(ctx, hipMemcpyAsync(dst + w0, pb, w1 - w0, hipMemcpyHostToDevice, ctx->stream));
        HIPC(ctx, hipEventRecord(ctx->pin_ev[k], ctx->stream));
        ctx->pin_busy[k] = true;
    }
    return 0;
}

int digest_common(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t n, const uint8_t* d_base, const uint64_t* off,
                  const uint64_t* lens, const char* const* names, const uint32_t* dd, int32_t* status,
                  const uint8_t* const* host_bufs, IndexCall* ic = nullptr)
{
    if (!ctx || !set || (n && (!lens || !status)))
        return fail(ctx, "null argument");
    for (uint32_t i = 0; i < n; i++) {
        if (dd && dd[i] > SDBFX_TILE)
            return fail(ctx, "dd block size above 16384");
        if (dd && dd[i] && dd[i] < SDBFX_MIN_FILE)
            return fail(ctx, "dd block size below 512");
        if (dd && dd[i] % 16)
            return fail(ctx, "dd block size must be a multiple of 16");
        status[i] = lens[i] < SDBFX_MIN_FILE ? -3 : 0;
    }
    // a failing call leaves the set as it was (inputs are digested in passes
    // of up to kPassBytes, each appending its digests)
    const size_t nd0 = set->names.size();
    const uint64_t nf0 = set->nfilt;
    auto rollback = [&](int rc) {
        set->names.resize(nd0);
        set->fold.resize(nd0);
        set->orig.resize(nd0);
        set->first.resize(nd0);
        set->bfc.resize(nd0);
        set->last.resize(nd0);
        set->dd.resize(nd0);
        set->maxe.resize(nd0);
        set->last_ec.resize(nd0);
        set->nfilt = nf0;
        return rc;
    };
    uint32_t i0 = 0;
    std::vector<uint64_t> hoff;
    while (i0 < n) {
        uint32_t i1 = i0;
        uint64_t bytes = 0;
        // with an index, a pass holds inputs of one mode only, so that the
        // index sees stream and block inputs in the order given
        auto blk = [&](uint32_t i) { return dd && dd[i]; };
        const uint64_t pass_bytes = host_bufs ? kPassBytes : kPassBytesDevice;
        while (i1 < n && (i1 == i0 || bytes + lens[i1] <= pass_bytes) &&
               !(ic && ic->ix && i1 > i0 && blk(i1) != blk(i0))) {
            bytes += (lens[i1] + SDBFX_PAD - 1) / SDBFX_PAD * SDBFX_PAD;
            i1++;
        }
        const uint8_t* base = d_base;
        const uint64_t* o = off;
        if (host_bufs) {
            // stage this pass's inputs into padded device memory
            hoff.assign(n, 0);
            uint64_t pos = SDBFX_PAD;
            for (uint32_t i = i0; i < i1; i++) {
                hoff[i] = pos;
                pos += (lens[i] + SDBFX_PAD - 1) / SDBFX_PAD * SDBFX_PAD;
            }
            if (int rc = ensure(ctx, ctx->stage, pos + SDBFX_PAD))
                return rollback(rc);
            if (int rc = upload_pass(ctx, (uint8_t*)ctx->stage.p, i0, i1, hoff, host_bufs, lens, pos + SDBFX_PAD))
                return rollback(rc);
            base = (const uint8_t*)ctx->stage.p;
            o = hoff.data();
        }
        if (int rc = digest_pass(ctx, set, i0, i1, base, o, lens, names, dd, status, ic))
            return rollback(rc);
        i0 = i1;
    }
    return 0;
}

// elem count of the last filter (sdbf::get_elem_count(s, bf_count-1))
struct PairRef {
    uint32_t set;     // 0 = A, 1 = B
    uint32_t idx;
    uint32_t count;   // filters taken from the digest (all, or a sample's first ones)
};

// Kernel choice for SDBFX_CMP_AUTO from the padded work of each tiling, per
// target digest with I reference-filter instances and T target filters:
// VALU ceil(I/64)*64 * T pair slots at about 8 SIMD cycles each (128
// instructions per lane-pair), MFMA (FP4 form) ceil(I/128)*128 *
// ceil(T/128)*128 slots at about 2 cycles each (measured on MI355X,
// candidate re-checks included).
bool use_mfma(const sdbfx_ctx* ctx, const std::vector<uint64_t>& inst, const std::vector<uint64_t>& tgt)
{
    // the MFMA kernel lists a candidate's target filter (its index in the
    // digest) in 25 bits
    for (uint64_t t : tgt)
        if (t >= (1ull << 25))
            return false;
    if (ctx->cmp_method != SDBFX_CMP_AUTO)
        return ctx->cmp_method == SDBFX_CMP_MFMA || ctx->cmp_method == SDBFX_CMP_MFMA_I8;
    double valu = 0, mfma = 0;
    for (size_t g = 0; g < inst.size(); g++) {
        valu += 8.0 * (double)((inst[g] + 63) / 64 * 64) * (double)tgt[g];
        mfma += 2.0 * (double)((inst[g] + 127) / 128 * 128) * (double)((tgt[g] + 127) / 128 * 128);
    }
    return mfma < valu;
}

// whether the MFMA kernel can take these target digests: it lists a
// candidate's target filter (its index in the digest) in 25 bits
bool use_mfma_possible(const std::vector<uint64_t>& tgt)
{
    for (uint64_t t : tgt)
        if (t >= (1ull << 25))
            return false;
    return true;
}

int launch_pairs(sdbfx_ctx* ctx, const sdbfx_set* const S[2], bool all, uint32_t na, uint64_t npairs,
                 const std::vector<PairRef>& ref, const std::vector<uint32_t>& order, const std::vector<uint32_t>& kc,
                 uint32_t nkeys, int mode, bool mfma, uint32_t chunk, int32_t* sorted);

// Scoring mode of one oriented pair.  sdbf_max_score takes the cutoff table
// and bit count from the reference digest's fastmode flag only
// (sdbf/sdbf_core.cc:422-437), so a pair whose reference digest is folded
// is scored on 64 bytes (CMP_MODE_FAST) whatever its target, and a plain
// reference against a folded target is scored on all 256 bytes with the
// target's folded Hamming weight (CMP_MODE_MIXED): there the denominator
// max_est - cut_off is an unsigned 32-bit difference that can wrap or be 0,
// and the fraction can exceed 1.
enum { CMP_MODE_PLAIN = 0, CMP_MODE_FAST = 1, CMP_MODE_MIXED = 2 };

// One comparison launch sequence over the oriented pairs `sel` (indices into
// ref / tgt / out), all of one scoring mode.
int compare_pairs(sdbfx_ctx* ctx, const sdbfx_set* const S[2], bool all, uint32_t na, uint32_t nb,
                  const std::vector<PairRef>& ref, const std::vector<PairRef>& tgt, const std::vector<uint32_t>& sel,
                  int mode, int32_t* out)
{
    const uint64_t npairs = sel.size();
    // group pairs by target digest (counting sort on a dense key)
    const uint32_t nkeys = na + (all ? 0 : nb);
    auto key = [&](const PairRef& r) { return r.set ? na + r.idx : r.idx; };
    std::vector<uint32_t> kc(nkeys + 1, 0);
    for (uint64_t s = 0; s < npairs; s++)
        kc[key(tgt[sel[s]]) + 1]++;
    for (uint32_t q = 0; q < nkeys; q++)
        kc[q + 1] += kc[q];
    std::vector<uint32_t> order(npairs);
    {
        std::vector<uint32_t> pos(kc.begin(), kc.end() - 1);
        for (uint64_t s = 0; s < npairs; s++)
            order[pos[key(tgt[sel[s]])]++] = sel[s];
    }
    std::vector<uint64_t> ginst, gtgt;
    for (uint32_t q = 0; q < nkeys; q++) {
        if (kc[q] == kc[q + 1])
            continue;
        const sdbfx_set* T = S[q >= na ? 1 : 0];
        uint64_t ic = 0;
        for (uint32_t s = kc[q]; s < kc[q + 1]; s++) {
            const PairRef r = ref[order[s]];
            ic += r.count;
        }
        ginst.push_back(ic);
        gtgt.push_back(T->bfc[q >= na ? q - na : q]);
    }
    const bool mfma_ok = mode == CMP_MODE_PLAIN && use_mfma_possible(gtgt);
    bool mfma = mode == CMP_MODE_PLAIN && use_mfma(ctx, ginst, gtgt);
    // target filters per work item.  The MFMA kernel's fixed costs (row
    // staging, publishing maxima) are per work item and its pruning works
    // across the tiles of one, so by default it takes the largest chunk that
    // still leaves a work item for every one of its 1024 wave slots.
    uint32_t chunk = SDBFX_CMP_CHUNK;
    if (mfma) {
        for (uint32_t ch = 16384; ch > SDBFX_CMP_CHUNK; ch >>= 1) {
            uint64_t n = 0;
            for (size_t g = 0; g < ginst.size(); g++)
                n += (ginst[g] + SDBFX_MROWS - 1) / SDBFX_MROWS * ((gtgt[g] + ch - 1) / ch);
            if (n >= 1024) {
                chunk = ch;
                break;
            }
        }
    }
    // Autotuning: for a shape class not seen before, time the VALU kernel
    // and the MFMA kernel at every chunk size on this call, twice each (all give
    // the same scores, which is checked), and keep the fastest for the class.
    if (ctx->autotune && ctx->cmp_method == SDBFX_CMP_AUTO && mode == CMP_MODE_PLAIN && mfma_ok) {
        uint64_t itot = 0, tmax = 0;
        for (size_t g = 0; g < ginst.size(); g++) {
            itot += ginst[g];
            tmax = std::max<uint64_t>(tmax, gtgt[g]);
        }
        auto lg = [](uint64_t v) { return (uint64_t)(v ? 64 - __builtin_clzll(v) : 0); };
        const uint64_t key = lg(itot) << 16 | lg(tmax) << 8 | lg(ginst.size());
        auto it = ctx->tuned.find(key);
        if (it == ctx->tuned.end()) {
            const uint32_t cands[6] = {0, 1024, 2048, 4096, 8192, 16384};   // 0: the VALU kernel
            sdbfx_tune_entry best{};
            best.ms = 1e30;
            std::vector<int32_t> keep(npairs), tmp(npairs);
            for (int k = 0; k < 6; k++) {
                // the better of two runs of each configuration
                float ms = 1e30f;
                for (int rep = 0; rep < 2; rep++) {
                    hipEvent_t e0 = get_event(ctx), e1 = get_event(ctx);
                    if (!e0 || !e1)
                        return fail(ctx, "hipEventCreate failed");
                    HIPC(ctx, hipEventRecord(e0, ctx->stream));
                    const int rc = launch_pairs(ctx, S, all, na, npairs, ref, order, kc, nkeys, mode, cands[k] != 0,
                                                cands[k] ? cands[k] : SDBFX_CMP_CHUNK, tmp.data());
                    if (rc != 0)
                        return rc;
                    HIPC(ctx, hipEventRecord(e1, ctx->stream));
                    HIPC(ctx, hipEventSynchronize(e1));
                    float m = 0.f;
                    HIPC(ctx, hipEventElapsedTime(&m, e0, e1));
                    ctx->evpool.push_back(e0);
                    ctx->evpool.push_back(e1);
                    ms = std::min(ms, m);
                }
                if (k == 0)
                    keep = tmp;
                else if (tmp != keep)
                    return fail(ctx, "autotuning: the comparison configurations disagree");
                if (ms < best.ms) {
                    best.ms = ms;
                    best.method = cands[k] ? SDBFX_CMP_MFMA : SDBFX_CMP_VALU;
                    best.chunk = cands[k] ? cands[k] : SDBFX_CMP_CHUNK;
                }
            }
            best.instances = itot;
            best.max_target = tmax;
            best.target_digests = (uint32_t)ginst.size();
            ctx->tuned[key] = best;
            for (uint64_t q = 0; q < npairs; q++)
                out[order[q]] = keep[q];
            return 0;
        }
        mfma = it->second.method == SDBFX_CMP_MFMA;
        chunk = it->second.chunk;
    }
    std::vector<int32_t> sorted(npairs);
    if (int rc = launch_pairs(ctx, S, all, na, npairs, ref, order, kc, nkeys, mode, mfma, chunk, sorted.data()))
        return rc;
    for (uint64_t q = 0; q < npairs; q++)
        out[order[q]] = sorted[q];
    return 0;
}

// One comparison launch sequence with a given kernel and chunk size over the
// pairs grouped by target digest (order / kc from compare_pairs); the scores
// in grouped order go to sorted.
int launch_pairs(sdbfx_ctx* ctx, const sdbfx_set* const S[2], bool all, uint32_t na, uint64_t npairs,
                 const std::vector<PairRef>& ref, const std::vector<uint32_t>& order, const std::vector<uint32_t>& kc,
                 uint32_t nkeys, int mode, bool mfma, uint32_t chunk, int32_t* sorted)
{
    const uint32_t rows = mfma ? SDBFX_MROWS : SDBFX_CMP_ROWS;
    std::vector<CmpPair> pairs(npairs);
    std::vector<CmpGroup> groups;
    uint64_t inst = 0, wi = 0, partial_n = 0;
    for (uint32_t q = 0; q < nkeys; q++) {
        if (kc[q] == kc[q + 1])
            continue;
        const PairRef t{q >= na ? 1u : 0u, q >= na ? q - na : q};
        const sdbfx_set* T = S[t.set];
        CmpGroup g{};
        g.tgt_first = T->first[t.idx];
        g.tgt_count = T->bfc[t.idx];
        g.chunk = chunk;
        g.nch = (g.tgt_count + chunk - 1) / chunk;
        g.inst_begin = inst;
        g.pair_begin = kc[q];
        g.pair_end = kc[q + 1];
        g.tgt_set = t.set;
        g.partial_begin = partial_n;
        for (uint32_t s = kc[q]; s < kc[q + 1]; s++) {
            const PairRef r = ref[order[s]];
            const sdbfx_set* R = S[r.set];
            CmpPair& p = pairs[s];
            p.ref_first = R->first[r.idx];
            p.ref_count = r.count;
            p.nch = g.nch;
            p.inst_begin = inst;
            p.partial_begin = g.partial_begin + (inst - g.inst_begin) * g.nch;
            p.ref_set = r.set;
            p.pad = 0;
            inst += p.ref_count;
        }
        const uint64_t ic = inst - g.inst_begin;
        if (ic > 0xFFFFFFFFull)
            return fail(ctx, "too many filter instances for one target digest");
        g.inst_count = (uint32_t)ic;
        g.nblk = (g.inst_count + rows - 1) / rows;
        g.wi_begin = wi;
        wi += (uint64_t)g.nblk * g.nch;
        partial_n += (uint64_t)g.inst_count * g.nch;
        groups.push_back(g);
    }

    int rc;
    if ((rc = upload(ctx, ctx->groups, groups)) || (rc = upload(ctx, ctx->pairs, pairs)) ||
        (rc = ensure(ctx, ctx->partial, partial_n * (mode == CMP_MODE_MIXED ? 8 : 4) + 256)) ||
        (rc = ensure(ctx, ctx->imax, inst * (mode == CMP_MODE_MIXED ? 8 : 4) + 256)) ||
        (rc = ensure(ctx, ctx->scores, npairs * 4 + 256)))
        return rc;
    if (wi && mfma && ((rc = ensure(ctx, ctx->gbest, inst * 8 + 256)) || (rc = ensure(ctx, ctx->rmap, inst * 8 + 256)) ||
                       (rc = ensure(ctx, ctx->rcode, inst * 4 + 256))))
        return rc;
    auto view = [](const sdbfx_set* X) {
        return SetView{(const uint32_t*)X->filt.p, (const uint16_t*)X->ecf.p, (const uint16_t*)X->hamf.p,
                       (const uint32_t*)X->cord.p, (const uint32_t*)X->ckey.p};
    };
    const SetView va = view(S[0]), vb = view(S[1]);
    if (wi && mfma) {
        // the column order of every target digest not ordered yet (after
        // any earlier build, possibly on another stream).  The sets stay
        // locked until the builds are queued and their events recorded.
        const int nsets = all ? 1 : 2;
        std::unique_lock<std::mutex> lk0(S[0]->cmu, std::defer_lock), lk1(S[1]->cmu, std::defer_lock);
        if (nsets == 2)
            std::lock(lk0, lk1);
        else
            lk0.lock();
        std::vector<uint64_t> cfirst[2];
        std::vector<uint32_t> ccount[2];
        int last = -1;   // the set whose build goes into the set-up launch
        for (int si = 0; si < nsets; si++) {
            const sdbfx_set* T = S[si];
            if (T->cev)
                HIPC(ctx, hipStreamWaitEvent(ctx->stream, T->cev, 0));
            T->cbuilt.resize(T->names.size(), 0);
            for (const CmpGroup& g : groups) {
                if ((int)(g.tgt_set && !all ? 1 : 0) != si)
                    continue;
                // the digest of this group: the one whose first filter it is
                const uint32_t idx = (uint32_t)(std::lower_bound(T->first.begin(), T->first.end(), g.tgt_first) -
                                                T->first.begin());
                if (idx < T->first.size() && !T->cbuilt[idx]) {
                    T->cbuilt[idx] = 1;
                    cfirst[si].push_back(g.tgt_first);
                    ccount[si].push_back(g.tgt_count);
                }
            }
            if (!cfirst[si].empty())
                last = si;
        }
        auto record = [&](const sdbfx_set* T) {
            if (!T->cev)
                HIPC(ctx, hipEventCreate(&T->cev));
            HIPC(ctx, hipEventRecord(T->cev, ctx->stream));
            return 0;
        };
        for (int si = 0; si < nsets; si++) {
            if (cfirst[si].empty() || si == last)
                continue;
            const sdbfx_set* T = S[si];
            if ((rc = upload(ctx, ctx->cofirst, cfirst[si])) || (rc = upload(ctx, ctx->cocount, ccount[si])))
                return rc;
            LAUNCH(ctx, col_order_kernel, (uint32_t)cfirst[si].size(), 1024, (const uint16_t*)T->ecf.p,
                   (const uint16_t*)T->hamf.p, (const uint64_t*)ctx->cofirst.p, (const uint32_t*)ctx->cocount.p,
                   (uint32_t*)T->cord.p, (uint32_t*)T->ckey.p);
            if ((rc = record(T)))
                return rc;
        }
        // one launch: the last column-order build, and per instance its row
        // data and its shared maximum for pruning cleared
        const sdbfx_set* T = S[last >= 0 ? last : 0];
        const uint32_t nd = last >= 0 ? (uint32_t)cfirst[last].size() : 0;
        if (nd && ((rc = upload(ctx, ctx->cofirst, cfirst[last])) || (rc = upload(ctx, ctx->cocount, ccount[last]))))
            return rc;
        LAUNCH(ctx, cmp_prep_kernel, nd + (uint32_t)((inst + 1023) / 1024), 1024, (const uint16_t*)T->ecf.p,
               (const uint16_t*)T->hamf.p, (const uint64_t*)ctx->cofirst.p, (const uint32_t*)ctx->cocount.p, nd,
               (uint32_t*)T->cord.p, (uint32_t*)T->ckey.p, va, vb, (const CmpPair*)ctx->pairs.p, (uint32_t)npairs,
               inst, (uint64_t*)ctx->rmap.p, (uint32_t*)ctx->rcode.p, (unsigned long long*)ctx->gbest.p);
        if (nd && (rc = record(T)))
            return rc;
        // one wave per SIMD (the 128 x 128 tile holds 256 accumulators per
        // lane): a 4-wave workgroup per CU, a multiple of the 8 XCDs
        const uint32_t grid = (uint32_t)std::min<uint64_t>(((wi + 3) / 4 + 7) & ~7ull, 256);
        if (ctx->cmp_method == SDBFX_CMP_MFMA_I8)
            LAUNCH(ctx, cmp_mfma_i8_kernel, grid, 256, va, vb, (const CmpGroup*)ctx->groups.p,
                   (uint32_t)groups.size(), (const uint64_t*)ctx->rmap.p, (const uint32_t*)ctx->rcode.p, wi,
                   (uint32_t*)ctx->partial.p, (unsigned long long*)ctx->gbest.p);
        else
            LAUNCH(ctx, cmp_mfma_kernel, grid, 256, va, vb, (const CmpGroup*)ctx->groups.p,
                   (uint32_t)groups.size(), (const uint64_t*)ctx->rmap.p, (const uint32_t*)ctx->rcode.p, wi,
                   (uint32_t*)ctx->partial.p, (unsigned long long*)ctx->gbest.p);
    } else if (wi) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(wi, 16384);
        if (mode == CMP_MODE_FAST)
            LAUNCH(ctx, cmp_fast_kernel, grid, 64, va, vb, (const CmpGroup*)ctx->groups.p,
                   (uint32_t)groups.size(), (const CmpPair*)ctx->pairs.p, wi, ctx->partial.p);
        else if (mode == CMP_MODE_MIXED)
            LAUNCH(ctx, cmp_mixed_kernel, grid, 64, va, vb, (const CmpGroup*)ctx->groups.p,
                   (uint32_t)groups.size(), (const CmpPair*)ctx->pairs.p, wi, ctx->partial.p);
        else
            LAUNCH(ctx, cmp_kernel, grid, 64, va, vb, (const CmpGroup*)ctx->groups.p, (uint32_t)groups.size(),
                   (const CmpPair*)ctx->pairs.p, wi, ctx->partial.p);
    }
    if (mode == CMP_MODE_MIXED) {
        if (inst)
            LAUNCH(ctx, chunk_max_wide_kernel, (uint32_t)((inst + 255) / 256), 256, (const CmpPair*)ctx->pairs.p,
                   (uint32_t)npairs, inst, (const uint64_t*)ctx->partial.p, (uint64_t*)ctx->imax.p);
        LAUNCH(ctx, pair_sum_wide_kernel, (uint32_t)((npairs + 255) / 256), 256, (const CmpPair*)ctx->pairs.p,
               (uint32_t)npairs, va.ecf, vb.ecf, (const uint64_t*)ctx->imax.p, (int32_t*)ctx->scores.p);
    } else {
        if (inst)
            LAUNCH(ctx, chunk_max_kernel, (uint32_t)((inst + 255) / 256), 256, (const CmpPair*)ctx->pairs.p,
                   (uint32_t)npairs, inst, (const uint32_t*)ctx->partial.p, (uint32_t*)ctx->imax.p);
        LAUNCH(ctx, pair_sum_kernel, (uint32_t)((npairs + 255) / 256), 256, (const CmpPair*)ctx->pairs.p,
               (uint32_t)npairs, va.ecf, vb.ecf, (const uint32_t*)ctx->imax.p, (int32_t*)ctx->scores.p);
    }
    if ((rc = download(ctx, sorted, ctx->scores.p, npairs * 4, ctx->scores.cap)))
        return rc;
    for (uint64_t s = 0; s < npairs; s++)
        if (sorted[s] == SDBFX_SCORE_NONFINITE)
            return fail(ctx, "a pair's score is not finite (a plain reference filter against a folded target "
                             "filter whose Hamming weight equals the cutoff): the reference aborts on it "
                             "(boost::math::round of infinity)");
    return 0;
}

// digests [a0, a1) of A against B (or, with `all`, every pair i < j of A)
int compare_impl(sdbfx_ctx* ctx, const sdbfx_set* A, const sdbfx_set* B, bool all, uint32_t sample, uint32_t a0,
                 uint32_t a1, int32_t* out)
{
    const uint32_t na = (uint32_t)A->names.size();
    const uint32_t nb = all ? na : (uint32_t)B->names.size();
    const uint64_t npairs = all ? (uint64_t)na * (na ? na - 1 : 0) / 2 : (uint64_t)(a1 - a0) * nb;
    if (npairs == 0)
        return 0;
    if (npairs > 0xFFFFFFFFull)
        return fail(ctx, "too many digest pairs in one call");
    const sdbfx_set* const S[2] = {A, all ? A : B};

    // Sampling (sdbf_score's `sample`, sdbf_core.cc:342-347): a calling
    // digest with more than `sample` filters takes part with its first
    // `sample` ones, and the swap rule sees that count and the element count
    // of filter sample - 1.  (The reference's random offsets, :370-372, are
    // never drawn: after the swap rule the reference side never has more than
    // `sample` filters.)  The calling digest is the row digest x below.
    std::vector<uint16_t> ecA;   // element counts of A's block-mode digests, when sampled
    auto ec_at = [&](uint32_t idx, uint32_t f) -> uint32_t {
        if (!A->dd[idx])
            return f + 1 < A->bfc[idx] ? A->maxe[idx] : A->last[idx];
        return ecA[A->first[idx] + f];
    };
    if (sample) {
        bool need = false;
        for (uint32_t i = a0; i < a1; i++)
            need |= A->dd[i] && A->bfc[i] > sample;
        if (need) {
            ecA.resize(A->nfilt);
            if (int rc = download(ctx, ecA.data(), A->ecf.p, A->nfilt * 2, A->ecf.cap))
                return rc;
        }
    }
    // orientation of every pair, in output order (sdbf_core.cc:349-359)
    std::vector<PairRef> ref(npairs), tgt(npairs);
    std::vector<uint32_t> sel[3];
    {
        uint64_t k = 0;
        for (uint32_t i = a0; i < a1; i++) {
            const uint32_t j0 = all ? i + 1 : 0;
            const bool sampled = sample && A->bfc[i] > sample;
            const uint32_t n1 = sampled ? sample : A->bfc[i];
            const uint32_t e1 = sampled ? ec_at(i, n1 - 1) : A->last_ec[i];
            for (uint32_t j = j0; j < nb; j++, k++) {
                PairRef x{0, i, n1}, y{all ? 0u : 1u, j, 0};
                const sdbfx_set* X = S[x.set];
                const sdbfx_set* Y = S[y.set];
                const uint32_t n2 = Y->bfc[y.idx];
                y.count = n2;
                const bool swap = n1 > n2 ||
                                  (n1 == n2 && (e1 > Y->last_ec[y.idx] ||
                                                strcmp(X->names[x.idx].c_str(), Y->names[y.idx].c_str()) > 0));
                if (swap)
                    x.count = X->bfc[x.idx];
                ref[k] = swap ? y : x;
                tgt[k] = swap ? x : y;
                const bool rf = S[ref[k].set]->fold[ref[k].idx] != 0, tf = S[tgt[k].set]->fold[tgt[k].idx] != 0;
                sel[rf ? CMP_MODE_FAST : tf ? CMP_MODE_MIXED : CMP_MODE_PLAIN].push_back((uint32_t)k);
            }
        }
    }
    for (int mode = 0; mode < 3; mode++)
        if (!sel[mode].empty())
            if (int rc = compare_pairs(ctx, S, all, na, nb, ref, tgt, sel[mode], mode, out))
                return rc;
    return 0;
}


}  // namespace

// ====================================================================== C ABI

extern "C" {

int sdbfx_ctx_create(int device, uint64_t mem_budget, sdbfx_ctx** out)
{
    if (!out)
        return -1;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return -1;
    if (hipSetDevice(device) != hipSuccess)
        return -1;
    if (scratch_setup(device) != 0)
        return -1;
    sdbfx_ctx* c = new sdbfx_ctx();
    c->device = device;
    c->budget = mem_budget;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return -1;
    }
    *out = c;
    return 0;
}

void sdbfx_ctx_destroy(sdbfx_ctx* ctx)
{
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (DevBuf* b : scratch_bufs(ctx))
        if (b->p)
            (void)hipFree(b->p);
    for (int k = 0; k < 2; k++) {
        if (ctx->pin[k])
            (void)hipHostFree(ctx->pin[k]);
        if (ctx->pin_ev[k])
            (void)hipEventDestroy(ctx->pin_ev[k]);
    }
    if (ctx->ring)
        (void)hipHostFree(ctx->ring);
    if (ctx->rb)
        (void)hipHostFree(ctx->rb);
    if (ctx->ring_ev)
        (void)hipEventDestroy(ctx->ring_ev);
    for (auto& ev : ctx->pending) {
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    for (hipEvent_t e : ctx->evpool)
        (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* sdbfx_last_error(const sdbfx_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int sdbfx_sync(sdbfx_ctx* ctx)
{
    if (!ctx)
        return -1;
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

uint64_t sdbfx_ctx_scratch_bytes(const sdbfx_ctx* ctx) { return ctx ? ctx->used : 0; }

int sdbfx_ctx_trim(sdbfx_ctx* ctx)
{
    if (!ctx)
        return -1;
    HIPC(ctx, hipSetDevice(ctx->device));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    for (DevBuf* b : scratch_bufs(ctx))
        if (b->p) {
            HIPC(ctx, hipFree(b->p));
            b->p = nullptr;
            b->cap = 0;
        }
    ctx->used = 0;
    return 0;
}

int sdbfx_set_create(sdbfx_ctx* ctx, sdbfx_set** out)
{
    if (!ctx || !out)
        return -1;
    sdbfx_set* s = new sdbfx_set();
    s->device = ctx->device;
    *out = s;
    return 0;
}

void sdbfx_set_destroy(sdbfx_set* set)
{
    if (!set)
        return;
    (void)hipSetDevice(set->device);
    (void)hipDeviceSynchronize();
    if (set->filt.p)
        (void)hipFree(set->filt.p);
    if (set->ecf.p)
        (void)hipFree(set->ecf.p);
    if (set->hamf.p)
        (void)hipFree(set->hamf.p);
    if (set->cord.p)
        (void)hipFree(set->cord.p);
    if (set->ckey.p)
        (void)hipFree(set->ckey.p);
    if (set->cev)
        (void)hipEventDestroy(set->cev);
    delete set;
}

uint32_t sdbfx_set_size(const sdbfx_set* set) { return set ? (uint32_t)set->names.size() : 0; }

uint64_t sdbfx_set_filter_count(const sdbfx_set* set) { return set ? set->nfilt : 0; }

int sdbfx_set_digest(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t n, const uint8_t* const* bufs,
                     const uint64_t* lens, const char* const* names, const uint32_t* dd_block,
                     int32_t* status)
{
    if (!ctx)
        return -1;
    if (n && !bufs)
        return fail(ctx, "null buffers");
    HIPC(ctx, hipSetDevice(ctx->device));
    return digest_common(ctx, set, n, nullptr, nullptr, lens, names, dd_block, status, bufs);
}

int sdbfx_set_digest_device(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t n, const void* d_base,
                            const uint64_t* offsets, const uint64_t* lens, const char* const* names,
                            const uint32_t* dd_block, int32_t* status)
{
    if (!ctx)
        return -1;
    if (n && (!d_base || !offsets))
        return fail(ctx, "null device input");
    for (uint32_t i = 0; i < n; i++)
        if (offsets[i] % SDBFX_PAD)
            return fail(ctx, "device input offsets must be multiples of 256");
    HIPC(ctx, hipSetDevice(ctx->device));
    return digest_common(ctx, set, n, (const uint8_t*)d_base, offsets, lens, names, dd_block, status,
                         nullptr);
}

// ---- indexes

int sdbfx_index_create(sdbfx_ctx* ctx, uint64_t size, sdbfx_index** out)
{
    if (!ctx || !out)
        return -1;
    *out = nullptr;
    if (size < 64 || (size & (size - 1)) || size > (256ull << 20))
        return fail(ctx, "index size must be a power of two in [64, 256 MiB]");
    HIPC(ctx, hipSetDevice(ctx->device));
    sdbfx_index* ix = new sdbfx_index;
    ix->device = ctx->device;
    ix->size = size;
    ix->mask = (uint32_t)(8 * size - 1);
    if (hipMalloc(&ix->bits, size) != hipSuccess || hipMalloc(&ix->elem, sizeof(unsigned long long)) != hipSuccess ||
        hipMemsetAsync(ix->bits, 0, size, ctx->stream) != hipSuccess ||
        hipMemsetAsync(ix->elem, 0, sizeof(unsigned long long), ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        sdbfx_index_destroy(ix);
        return fail(ctx, "index allocation failed");
    }
    *out = ix;
    return 0;
}

void sdbfx_index_destroy(sdbfx_index* ix)
{
    if (!ix)
        return;
    (void)hipSetDevice(ix->device);
    (void)hipDeviceSynchronize();
    if (ix->bits)
        (void)hipFree(ix->bits);
    if (ix->elem)
        (void)hipFree(ix->elem);
    if (ix->first)
        (void)hipFree(ix->first);
    delete ix;
}

uint64_t sdbfx_index_size(const sdbfx_index* ix) { return ix ? ix->size : 0; }

int sdbfx_index_read(sdbfx_ctx* ctx, const sdbfx_index* ix, uint8_t* bits, uint64_t* elem_count)
{
    if (!ctx || !ix)
        return -1;
    HIPC(ctx, hipSetDevice(ctx->device));
    if (bits)
        HIPC(ctx, hipMemcpyAsync(bits, ix->bits, ix->size, hipMemcpyDeviceToHost, ctx->stream));
    unsigned long long e = 0;
    HIPC(ctx, hipMemcpyAsync(&e, ix->elem, sizeof(e), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    if (elem_count)
        *elem_count = e;
    return 0;
}

int sdbfx_index_write(sdbfx_ctx* ctx, sdbfx_index* ix, const uint8_t* bits, uint64_t elem_count)
{
    if (!ctx || !ix || !bits)
        return -1;
    HIPC(ctx, hipSetDevice(ctx->device));
    const unsigned long long e = elem_count;
    HIPC(ctx, hipMemcpyAsync(ix->bits, bits, ix->size, hipMemcpyHostToDevice, ctx->stream));
    HIPC(ctx, hipMemcpyAsync(ix->elem, &e, sizeof(e), hipMemcpyHostToDevice, ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

namespace {

int digest_indexed(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t n, const uint8_t* d_base, const uint64_t* offsets,
                   const uint64_t* lens, const char* const* names, const uint32_t* dd_block, sdbfx_index* index,
                   const sdbfx_index* const* search, uint32_t nsearch, uint32_t* matches, uint64_t matches_cap,
                   int32_t* status, const uint8_t* const* bufs)
{
    if (nsearch && (!search || !matches))
        return fail(ctx, "null search indexes or matches");
    IndexCall ic;
    ic.ix = index;
    for (uint32_t m = 0; m < nsearch; m++) {
        if (!search[m] || search[m]->device != ctx->device)
            return fail(ctx, "searched index missing or on another device");
        ic.search.push_back(IndexRef{search[m]->bits, search[m]->mask, 0});
    }
    if (index && index->device != ctx->device)
        return fail(ctx, "index on another device");
    if (nsearch) {
        uint64_t need = 0;
        for (uint32_t i = 0; i < n; i++)
            if (dd_block && dd_block[i] && lens && lens[i] >= SDBFX_MIN_FILE)
                need += block_count(lens[i], dd_block[i]);
        if (need > matches_cap)
            return fail(ctx, "matches holds " + std::to_string(matches_cap) + " blocks, the call has " +
                                 std::to_string(need));
        ic.matches = matches;
    }
    HIPC(ctx, hipSetDevice(ctx->device));
    return digest_common(ctx, set, n, d_base, offsets, lens, names, dd_block, status, bufs,
                         (index || nsearch) ? &ic : nullptr);
}

}  // namespace

int sdbfx_set_digest_indexed(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t n, const uint8_t* const* bufs,
                             const uint64_t* lens, const char* const* names, const uint32_t* dd_block,
                             sdbfx_index* index, const sdbfx_index* const* search, uint32_t nsearch,
                             uint32_t* matches, uint64_t matches_cap, int32_t* status)
{
    if (!ctx)
        return -1;
    if (n && !bufs)
        return fail(ctx, "null buffers");
    return digest_indexed(ctx, set, n, nullptr, nullptr, lens, names, dd_block, index, search, nsearch, matches,
                          matches_cap, status, bufs);
}

int sdbfx_set_digest_device_indexed(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t n, const void* d_base,
                                    const uint64_t* offsets, const uint64_t* lens, const char* const* names,
                                    const uint32_t* dd_block, sdbfx_index* index, const sdbfx_index* const* search,
                                    uint32_t nsearch, uint32_t* matches, uint64_t matches_cap, int32_t* status)
{
    if (!ctx)
        return -1;
    if (n && (!d_base || !offsets))
        return fail(ctx, "null device input");
    for (uint32_t i = 0; i < n; i++)
        if (offsets[i] % SDBFX_PAD)
            return fail(ctx, "device input offsets must be multiples of 256");
    return digest_indexed(ctx, set, n, (const uint8_t*)d_base, offsets, lens, names, dd_block, index, search,
                          nsearch, matches, matches_cap, status, nullptr);
}

int sdbfx_set_add(sdbfx_ctx* ctx, sdbfx_set* set, const char* name, uint64_t orig_size, uint32_t bf_count,
                  uint32_t last_count, uint32_t dd_block_size, uint32_t max_elem, const uint8_t* filters,
                  const uint16_t* elem_counts)
{
    if (!ctx || !set || !filters)
        return fail(ctx, "null argument");
    if (bf_count == 0)
        return fail(ctx, "digest without filters");
    if (dd_block_size && !elem_counts)
        return fail(ctx, "block-mode digest needs element counts");
    HIPC(ctx, hipSetDevice(ctx->device));
    if (int rc = set_reserve(ctx, set, set->nfilt + bf_count))
        return rc;
    std::vector<uint16_t> ec(bf_count), ham(bf_count);
    for (uint32_t i = 0; i < bf_count; i++) {
        if (dd_block_size)
            ec[i] = elem_counts[i];
        else
            ec[i] = (uint16_t)(i + 1 < bf_count ? max_elem : last_count);
        uint32_t h = 0;
        for (int b = 0; b < SDBFX_BF; b++)
            h += (uint32_t)__builtin_popcount(filters[(uint64_t)i * SDBFX_BF + b]);
        ham[i] = (uint16_t)h;
    }
    HIPC(ctx, hipMemcpyAsync((uint8_t*)set->filt.p + set->nfilt * SDBFX_BF, filters,
                             (uint64_t)bf_count * SDBFX_BF, hipMemcpyHostToDevice, ctx->stream));
    HIPC(ctx, hipMemcpyAsync((uint16_t*)set->ecf.p + set->nfilt, ec.data(), bf_count * 2,
                             hipMemcpyHostToDevice, ctx->stream));
    HIPC(ctx, hipMemcpyAsync((uint16_t*)set->hamf.p + set->nfilt, ham.data(), bf_count * 2,
                             hipMemcpyHostToDevice, ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    set->names.push_back(name ? name : "");
    set->fold.push_back(0);
    set->orig.push_back(orig_size);
    set->first.push_back(set->nfilt);
    set->bfc.push_back(bf_count);
    set->last.push_back(dd_block_size ? 0 : last_count);
    set->dd.push_back(dd_block_size);
    set->maxe.push_back(max_elem);
    set->last_ec.push_back(ec[bf_count - 1]);
    set->nfilt += bf_count;
    return 0;
}

int sdbfx_set_info(const sdbfx_set* set, uint32_t idx, sdbfx_digest_info* info)
{
    if (!set || !info || idx >= set->names.size())
        return -1;
    info->name = set->names[idx].c_str();
    info->orig_size = set->orig[idx];
    info->bf_count = set->bfc[idx];
    info->last_count = set->last[idx];
    info->dd_block_size = set->dd[idx];
    info->max_elem = set->maxe[idx];
    return 0;
}

int sdbfx_set_read(sdbfx_ctx* ctx, const sdbfx_set* set, uint32_t idx, uint8_t* filters,
                   uint16_t* elem_counts, uint16_t* hamming)
{
    if (!ctx || !set || idx >= set->names.size())
        return fail(ctx, "bad digest index");
    HIPC(ctx, hipSetDevice(ctx->device));
    const uint64_t f = set->first[idx], c = set->bfc[idx];
    if (filters)
        HIPC(ctx, hipMemcpyAsync(filters, (const uint8_t*)set->filt.p + f * SDBFX_BF, c * SDBFX_BF,
                                 hipMemcpyDeviceToHost, ctx->stream));
    if (elem_counts)
        HIPC(ctx, hipMemcpyAsync(elem_counts, (const uint16_t*)set->ecf.p + f, c * 2, hipMemcpyDeviceToHost,
                                 ctx->stream));
    if (hamming)
        HIPC(ctx, hipMemcpyAsync(hamming, (const uint16_t*)set->hamf.p + f, c * 2, hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int sdbfx_set_to_text(sdbfx_ctx* ctx, const sdbfx_set* set, char* out, uint64_t cap, uint64_t* len)
{
    if (!ctx || !set || !len)
        return fail(ctx, "null argument");
    HIPC(ctx, hipSetDevice(ctx->device));
    // headers and layout (sdbf/sdbf_class.cc:294-334)
    const uint32_t n = (uint32_t)set->names.size();
    std::string hdr;
    std::vector<TextDigest> td(n);
    uint64_t pos = 0, q = 0;
    for (uint32_t d = 0; d < n; d++) {
        char buf[96];
        std::string h = set->dd[d] ? "sdbf-dd:03:" : "sdbf:03:";
        h += std::to_string(set->names[d].size()) + ":" + set->names[d] + ":";
        if (set->dd[d])
            snprintf(buf, sizeof buf, "%llu:sha1:256:5:7ff:%u:%u:%u", (unsigned long long)set->orig[d],
                     set->maxe[d], set->bfc[d], set->dd[d]);
        else
            snprintf(buf, sizeof buf, "%llu:sha1:256:5:7ff:%u:%u:%u:", (unsigned long long)set->orig[d],
                     set->maxe[d], set->bfc[d], set->last[d]);
        h += buf;
        if (hdr.size() + h.size() > 0xFFFFFFFFull)
            return fail(ctx, "text headers too long");
        TextDigest& t = td[d];
        t.filt_first = set->first[d];
        t.q_begin = q;
        t.text_off = pos;
        t.body_off = pos + h.size();
        t.hdr_off = (uint32_t)hdr.size();
        t.hdr_len = (uint32_t)h.size();
        t.bf_count = set->bfc[d];
        t.dd = set->dd[d];
        const uint64_t raw = (uint64_t)t.bf_count * SDBFX_BF;
        const uint64_t quanta = t.dd ? (uint64_t)t.bf_count * 87 : (raw + 2) / 3;
        const uint64_t body = t.dd ? (uint64_t)t.bf_count * 348 : (raw + 2) / 3 * 4;
        q += quanta;
        pos = t.body_off + body + 1;   // + newline
        hdr += h;
    }
    *len = pos;
    if (!out)
        return 0;
    if (cap < pos)
        return fail(ctx, "text buffer too small");
    if (n == 0)
        return 0;
    int rc;
    if ((rc = ensure(ctx, ctx->text, pos + 256)) || (rc = upload(ctx, ctx->tdesc, td)) ||
        (rc = ensure(ctx, ctx->thdr, hdr.size() + 256)))
        return rc;
    HIPC(ctx, hipMemcpyAsync(ctx->thdr.p, hdr.data(), hdr.size(), hipMemcpyHostToDevice, ctx->stream));
    LAUNCH(ctx, text_header_kernel, n, 256, (const TextDigest*)ctx->tdesc.p, (const char*)ctx->thdr.p, pos,
           (char*)ctx->text.p);
    if (q)
        LAUNCH(ctx, text_b64_kernel, (uint32_t)((q + 255) / 256), 256, (const uint8_t*)set->filt.p,
               (const uint16_t*)set->ecf.p, (const TextDigest*)ctx->tdesc.p, n, q, (char*)ctx->text.p);
    HIPC(ctx, hipMemcpyAsync(out, ctx->text.p, pos, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int sdbfx_set_add_text(sdbfx_ctx* ctx, sdbfx_set* set, const char* text, uint64_t len, uint32_t* added)
{
    if (!ctx || !set || (!text && len))
        return fail(ctx, "null argument");
    if (added)
        *added = 0;
    HIPC(ctx, hipSetDevice(ctx->device));
    auto bad = [&](const char* why) {
        ctx->err = std::string("text: ") + why;
        return -2;   // the reference's "invalid format" (sdbf_class.cc:152)
    };
    struct Hdr {
        std::string name;
        uint64_t orig, body;
        uint32_t maxe, bfc, last, dd;
        uint16_t last_ec;
    };
    std::vector<Hdr> hs;
    uint64_t pos = 0, used = 0, nf = 0;
    while (pos < len) {
        const char* nl = (const char*)memchr(text + pos, '\n', len - pos);
        const uint64_t end = nl ? (uint64_t)(nl - text) : len;
        if (end == pos)
            break;   // an empty line ends the set (sdbf_set.cc:57-58)
        if (end - pos < 344) {
            ctx->err = "text: line too short";
            return -1;   // sdbf_class.cc:145-146
        }
        // header fields, ':'-separated, as sdbf_class.cc:149-215 reads them
        uint64_t p = pos;
        auto field = [&](std::string& f) {
            const char* c = (const char*)memchr(text + p, ':', end - p);
            const uint64_t e = c ? (uint64_t)(c - text) : end;
            f.assign(text + p, e - p);
            p = c ? e + 1 : end;
            return c != nullptr;
        };
        auto number = [&](uint64_t& v, int base) {
            std::string f;
            if (!field(f) || f.empty())
                return false;
            char* e = nullptr;
            v = strtoull(f.c_str(), &e, base);
            return *e == 0;
        };
        Hdr h{};
        std::string magic, f;
        uint64_t v[7];
        if (!field(magic) || magic.compare(0, 4, "sdbf") != 0)
            return bad("not an sdbf line");
        const bool dd = magic == "sdbf-dd";
        if (!field(f) || !field(f) || !field(h.name) || !number(v[0], 10) || !field(f) || !number(v[1], 10) ||
            !number(v[2], 10) || !number(v[3], 16) || !number(v[4], 10) || !number(v[5], 10))
            return bad("header");
        h.orig = v[0];
        if (v[1] != SDBFX_BF || v[5] == 0 || v[5] > 0xFFFFFFFFull || v[4] > 0xFFFFull)
            return bad("filter size or count");
        h.maxe = (uint32_t)v[4];
        h.bfc = (uint32_t)v[5];
        const uint64_t raw = (uint64_t)h.bfc * SDBFX_BF;
        if (dd) {
            if (!number(v[6], 10) || v[6] == 0)
                return bad("dd block size");
            h.dd = (uint32_t)v[6];
            h.body = p - 1;   // each filter: ':' 2 hex digits ':' 344 characters
            if (text[h.body] != ':' || end - h.body < (uint64_t)h.bfc * 348)
                return bad("block-mode body");
            const uint64_t lf = h.body + (uint64_t)(h.bfc - 1) * 348;
            unsigned x = 0;
            if (sscanf(std::string(text + lf + 1, 2).c_str(), "%2x", &x) != 1)
                return bad("element count");
            h.last_ec = (uint16_t)x;
        } else {
            std::string lc;
            // last count, then the base64 run (sdbf_class.cc:198-214)
            const char* c = (const char*)memchr(text + p, ':', end - p);
            if (!c)
                return bad("last count");
            lc.assign(text + p, (uint64_t)(c - text) - p);
            char* e = nullptr;
            h.last = (uint32_t)strtoul(lc.c_str(), &e, 10);
            if (lc.empty() || *e)
                return bad("last count");
            h.body = (uint64_t)(c - text) + 1;
            if (end - h.body < (raw + 2) / 3 * 4)
                return bad("stream body");
            h.last_ec = (uint16_t)h.last;
        }
        nf += h.bfc;
        used = end;
        hs.push_back(std::move(h));
        pos = end + 1;
    }
    if (hs.empty())
        return 0;
    // descriptors; the text goes to the device through the pinned buffers
    std::vector<TextParse> tp(hs.size());
    uint64_t q = 0, fcur = set->nfilt;
    for (size_t d = 0; d < hs.size(); d++) {
        TextParse& t = tp[d];
        t.filt_first = fcur;
        t.q_begin = q;
        t.body_off = hs[d].body;
        t.bf_count = hs[d].bfc;
        t.dd = hs[d].dd;
        t.max_elem = hs[d].maxe;
        t.last_count = hs[d].last;
        q += t.dd ? (uint64_t)t.bf_count * 87 : ((uint64_t)t.bf_count * SDBFX_BF + 2) / 3;
        fcur += t.bf_count;
    }
    int rc;
    if ((rc = set_reserve(ctx, set, set->nfilt + nf)) || (rc = ensure(ctx, ctx->text, used + 256)) ||
        (rc = upload(ctx, ctx->tdesc, tp)) || (rc = ensure(ctx, ctx->thdr, 256)))
        return rc;
    if ((rc = upload_host(ctx, (uint8_t*)ctx->text.p, (const uint8_t*)text, used)))
        return rc;
    HIPC(ctx, hipMemsetAsync(ctx->thdr.p, 0, 4, ctx->stream));
    LAUNCH(ctx, text_decode_kernel, (uint32_t)((q + 255) / 256), 256, (const char*)ctx->text.p,
           (const TextParse*)ctx->tdesc.p, (uint32_t)tp.size(), q, (uint8_t*)set->filt.p, (uint16_t*)set->ecf.p,
           (uint32_t*)ctx->thdr.p);
    LAUNCH(ctx, text_finish_kernel, (uint32_t)((nf + 255) / 256), 256, (const TextParse*)ctx->tdesc.p,
           (uint32_t)tp.size(), set->nfilt, nf, (const uint8_t*)set->filt.p, (uint16_t*)set->ecf.p,
           (uint16_t*)set->hamf.p);
    uint32_t err = 0;
    HIPC(ctx, hipMemcpyAsync(&err, ctx->thdr.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    if (err)
        return bad("base64 or element counts");
    for (size_t d = 0; d < hs.size(); d++) {
        set->names.push_back(hs[d].name);
        set->fold.push_back(0);
        set->orig.push_back(hs[d].orig);
        set->first.push_back(tp[d].filt_first);
        set->bfc.push_back(hs[d].bfc);
        set->last.push_back(hs[d].dd ? 0 : hs[d].last);
        set->dd.push_back(hs[d].dd);
        set->maxe.push_back(hs[d].maxe);
        set->last_ec.push_back(hs[d].last_ec);
    }
    set->nfilt += nf;
    if (added)
        *added = (uint32_t)hs.size();
    return 0;
}

}  // extern "C"

namespace {

// Folds digests [d0, d1) of the set that are not folded yet (sdbf::fast,
// sdbf/sdbf_class.cc:426-439): runs of consecutive plain digests are folded
// by one launch each.
int fold_range(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t d0, uint32_t d1)
{
    HIPC(ctx, hipSetDevice(ctx->device));
    {
        std::lock_guard<std::mutex> lk(set->cmu);
        if (set->cev)
            HIPC(ctx, hipStreamWaitEvent(ctx->stream, set->cev, 0));
        // Hamming weights change: column orders built before are void
        set->cbuilt.resize(set->names.size(), 0);
        for (uint32_t d = d0; d < d1; d++)
            if (!set->fold[d])
                set->cbuilt[d] = 0;
    }
    for (uint32_t d = d0; d < d1;) {
        if (set->fold[d]) {
            d++;
            continue;
        }
        uint32_t e = d;
        while (e < d1 && !set->fold[e])
            e++;
        const uint64_t f0 = set->first[d];
        const uint64_t nf = (e < set->first.size() ? set->first[e] : set->nfilt) - f0;
        if (nf)
            LAUNCH(ctx, fold_kernel, (uint32_t)((nf + 255) / 256), 256, (uint32_t*)set->filt.p,
                   (uint16_t*)set->hamf.p, f0, nf);
        d = e;
    }
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t d = d0; d < d1; d++)
        set->fold[d] = 1;
    return 0;
}

}  // namespace

extern "C" {

int sdbfx_set_fold(sdbfx_ctx* ctx, sdbfx_set* set)
{
    if (!ctx || !set)
        return fail(ctx, "null argument");
    return fold_range(ctx, set, 0, (uint32_t)set->names.size());
}

int sdbfx_set_fold_digest(sdbfx_ctx* ctx, sdbfx_set* set, uint32_t idx)
{
    if (!ctx || !set)
        return fail(ctx, "null argument");
    if (idx >= set->names.size())
        return fail(ctx, "bad digest index");
    return fold_range(ctx, set, idx, idx + 1);
}

int sdbfx_set_folded(const sdbfx_set* set)
{
    if (!set || set->names.empty())
        return 0;
    for (uint8_t f : set->fold)
        if (!f)
            return 0;
    return 1;
}

int sdbfx_set_digest_folded(const sdbfx_set* set, uint32_t idx)
{
    return set && idx < set->fold.size() && set->fold[idx] ? 1 : 0;
}

int sdbfx_compare_all(sdbfx_ctx* ctx, const sdbfx_set* set, int32_t* scores)
{
    if (!ctx || !set || !scores)
        return fail(ctx, "null argument");
    HIPC(ctx, hipSetDevice(ctx->device));
    return compare_impl(ctx, set, set, true, 0, 0, (uint32_t)set->names.size(), scores);
}

int sdbfx_compare_to(sdbfx_ctx* ctx, const sdbfx_set* q, const sdbfx_set* t, int32_t* scores)
{
    if (!ctx || !q || !t || !scores)
        return fail(ctx, "null argument");
    HIPC(ctx, hipSetDevice(ctx->device));
    return compare_impl(ctx, q, t, false, 0, 0, (uint32_t)q->names.size(), scores);
}

int sdbfx_compare_to_sampled(sdbfx_ctx* ctx, const sdbfx_set* q, const sdbfx_set* t, uint32_t sample,
                             int32_t* scores)
{
    if (!ctx || !q || !t || !scores)
        return fail(ctx, "null argument");
    HIPC(ctx, hipSetDevice(ctx->device));
    return compare_impl(ctx, q, t, false, sample, 0, (uint32_t)q->names.size(), scores);
}

int sdbfx_compare_to_range(sdbfx_ctx* ctx, const sdbfx_set* q, uint32_t q_first, uint32_t q_count,
                           const sdbfx_set* t, int32_t* scores)
{
    if (!ctx || !q || !t || !scores)
        return fail(ctx, "null argument");
    if ((uint64_t)q_first + q_count > q->names.size())
        return fail(ctx, "query digest range out of the set");
    HIPC(ctx, hipSetDevice(ctx->device));
    return compare_impl(ctx, q, t, false, 0, q_first, q_first + q_count, scores);
}

int sdbfx_device_alloc(sdbfx_ctx* ctx, uint64_t bytes, void** dptr)
{
    if (!ctx || !dptr)
        return -1;
    HIPC(ctx, hipSetDevice(ctx->device));
    HIPC(ctx, hipMalloc(dptr, bytes + SDBFX_PAD));
    HIPC(ctx, hipMemset((uint8_t*)*dptr + bytes, 0, SDBFX_PAD));
    return 0;
}

int sdbfx_device_free(sdbfx_ctx* ctx, void* dptr)
{
    if (!ctx)
        return -1;
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    HIPC(ctx, hipFree(dptr));
    return 0;
}

int sdbfx_upload(sdbfx_ctx* ctx, void* dst, const void* src, uint64_t bytes)
{
    if (!ctx)
        return -1;
    HIPC(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int sdbfx_set_compare_method(sdbfx_ctx* ctx, int method)
{
    if (!ctx)
        return -1;
    if (method != SDBFX_CMP_AUTO && method != SDBFX_CMP_VALU && method != SDBFX_CMP_MFMA &&
        method != SDBFX_CMP_MFMA_I8)
        return fail(ctx, "unknown compare method");
    ctx->cmp_method = method;
    return 0;
}

int sdbfx_set_autotune(sdbfx_ctx* ctx, int on)
{
    if (!ctx)
        return -1;
    ctx->autotune = on != 0;
    return 0;
}

int sdbfx_autotune_table(const sdbfx_ctx* ctx, sdbfx_tune_entry* entries, uint32_t cap, uint32_t* n)
{
    if (!ctx || !n)
        return -1;
    uint32_t k = 0;
    for (const auto& kv : ctx->tuned) {
        if (entries && k < cap)
            entries[k] = kv.second;
        k++;
    }
    *n = k;
    return 0;
}

int sdbfx_set_profiling(sdbfx_ctx* ctx, int on)
{
    if (!ctx)
        return -1;
    ctx->prof = on != 0;
    return 0;
}

int sdbfx_kernel_time(sdbfx_ctx* ctx, const char* name, double* ms, uint64_t* launches)
{
    if (!ctx || !name)
        return -1;
    if (int rc = collect_times(ctx))
        return rc;
    auto it = ctx->times.find(name);
    if (ms)
        *ms = it == ctx->times.end() ? 0.0 : it->second.first;
    if (launches)
        *launches = it == ctx->times.end() ? 0 : it->second.second;
    return 0;
}

}  // extern "C"

namespace {

// The checks of sdbfx_digest_blocks_noalloc; its dd-block count in *nblk.
int noalloc_check(sdbfx_ctx* ctx, const void* d_input, uint64_t len, uint32_t dd_block, const void* d_filters,
                  const void* d_elem_counts, const void* d_hamming, uint64_t* nblk)
{
    if (!ctx || !d_input || !d_filters || !d_elem_counts || !d_hamming)
        return fail(ctx, "null argument");
    if ((uintptr_t)d_input % SDBFX_PAD)
        return fail(ctx, "the input must be 256-byte aligned");
    if (dd_block < SDBFX_MIN_FILE || dd_block > SDBFX_TILE || dd_block % 16)
        return fail(ctx, "dd block size must be a multiple of 16 in [512, 16384]");
    if (len < SDBFX_MIN_FILE) {
        ctx->err = "input too small";
        return -3;
    }
    *nblk = block_count(len, dd_block);
    if (*nblk > SDBFX_NOALLOC_BLOCKS)
        return fail(ctx, "input of " + std::to_string(*nblk) + " dd-blocks: the no-allocation digest takes at most " +
                             std::to_string(SDBFX_NOALLOC_BLOCKS));
    if (!scratch_ptr(ctx->device))
        return fail(ctx, "no device address for the scratch array");
    return 0;
}

// The launches of sdbfx_digest_blocks_noalloc: the work list into the fixed
// scratch array, then one workgroup per dd-block.  Nothing else is queued,
// so the sequence can be captured into a graph.
int noalloc_launch(sdbfx_ctx* ctx, const void* d_input, uint64_t len, uint32_t dd_block, uint64_t nblk,
                   uint8_t* d_filters, uint16_t* d_elem_counts, uint16_t* d_hamming)
{
    BlockDesc* bd = (BlockDesc*)scratch_ptr(ctx->device);
    hipLaunchKernelGGL(block_desc_kernel, dim3((uint32_t)((nblk + 255) / 256)), dim3(256), 0, ctx->stream, bd,
                       len, dd_block, (uint32_t)nblk);
    HIPC(ctx, hipGetLastError());
    hipLaunchKernelGGL(block_digest_kernel, dim3((uint32_t)nblk), dim3(SDBFX_NT), 0, ctx->stream,
                       (const uint8_t*)d_input, (const BlockDesc*)bd, d_filters, d_elem_counts, d_hamming,
                       (uint32_t*)nullptr);
    HIPC(ctx, hipGetLastError());
    return 0;
}

}  // namespace

struct sdbfx_graph {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int sdbfx_graph_blocks_create(sdbfx_ctx* ctx, const void* d_input, uint64_t len, uint32_t dd_block,
                              uint8_t* d_filters, uint16_t* d_elem_counts, uint16_t* d_hamming, uint32_t* nfilters,
                              sdbfx_graph** out)
{
    if (!out || !nfilters)
        return fail(ctx, "null argument");
    *out = nullptr;
    *nfilters = 0;
    uint64_t nblk = 0;
    if (int rc = noalloc_check(ctx, d_input, len, dd_block, d_filters, d_elem_counts, d_hamming, &nblk))
        return rc;
    HIPC(ctx, hipSetDevice(ctx->device));
    HIPC(ctx, hipStreamSynchronize(ctx->stream));
    sdbfx_graph* g = new sdbfx_graph;
    g->device = ctx->device;
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        delete g;
        return fail(ctx, "hipStreamBeginCapture failed");
    }
    const int rc = noalloc_launch(ctx, d_input, len, dd_block, nblk, d_filters, d_elem_counts, d_hamming);
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
    if (rc != 0 || e != hipSuccess || hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        sdbfx_graph_destroy(g);
        return rc != 0 ? rc : fail(ctx, "graph capture failed");
    }
    *nfilters = (uint32_t)nblk;
    *out = g;
    return 0;
}

int sdbfx_graph_launch(sdbfx_ctx* ctx, sdbfx_graph* g)
{
    if (!ctx || !g || g->device != ctx->device)
        return fail(ctx, "null graph or graph of another device");
    HIPC(ctx, hipGraphLaunch(g->exec, ctx->stream));
    return 0;
}

void sdbfx_graph_destroy(sdbfx_graph* g)
{
    if (!g)
        return;
    (void)hipSetDevice(g->device);
    if (g->exec)
        (void)hipGraphExecDestroy(g->exec);
    if (g->graph)
        (void)hipGraphDestroy(g->graph);
    delete g;
}

int sdbfx_digest_blocks_noalloc(sdbfx_ctx* ctx, const void* d_input, uint64_t len, uint32_t dd_block,
                                uint8_t* d_filters, uint16_t* d_elem_counts, uint16_t* d_hamming, uint32_t* nfilters)
{
    if (!nfilters)
        return fail(ctx, "null argument");
    *nfilters = 0;
    uint64_t nblk = 0;
    if (int rc = noalloc_check(ctx, d_input, len, dd_block, d_filters, d_elem_counts, d_hamming, &nblk))
        return rc;
    HIPC(ctx, hipSetDevice(ctx->device));
    if (int rc = noalloc_launch(ctx, d_input, len, dd_block, nblk, d_filters, d_elem_counts, d_hamming))
        return rc;
    *nfilters = (uint32_t)nblk;
    return 0;
}

int sdbfx_reset_kernel_times(sdbfx_ctx* ctx)
{
    if (!ctx)
        return -1;
    if (int rc = collect_times(ctx))
        return rc;
    ctx->times.clear();
    return 0;
}

}  // extern "C"

// gfx950 kernels of the sdbf digest and comparison path.
//
// Digest (one 256-thread workgroup per 16 KiB dd-block, or per 16 Ki-position
// tile of a stream-mode unit):
//   1. ranks   : rank of every 64-byte window = ENTR64_RANKS[entropy >> 10],
//                entropy = sum over byte values of E[count]; each of 128
//                lanes rolls a private byte histogram over its run of windows
//                (sdbf/sdbf_core.cc:22-45, sdbf/entr64.cc:28-71).
//   2. F       : for every window start i, the position the reference's full
//                rescan (sdbf_core.cc:69-78) settles on: first occurrence of the
//                smallest non-zero rank in [i, i+63], moved to the end of the
//                run of equal ranks that starts there.  Computed with wave-wide
//                suffix/prefix-min scans over 64-aligned blocks and run-start
//                ballots.  Packed with the rank as rank(10 bits) | F-i (6 bits).
//   3. scores  : the popularity state machine of sdbf_core.cc:51-89.  Its state
//                before step i is the remembered minimum position m (the rank
//                at m is implied), so 256 lanes each run one segment of steps
//                after a speculative SDBFX_WARM-step warm-up.  Every lane's start
//                state is checked against its predecessor's end state and
//                mismatching segments are undone and re-run until all agree
//                (fix-point).  The first segment of a stream tile that does not
//                start at position 0 is certified by running all 64 possible
//                incoming states through the warm-up: if they coalesce, the
//                state is exact whatever came before.
//   4. block mode: score histogram, threshold/allowance (sdbf_core.cc:271-276),
//                candidate SHA-1 in parallel, then the ordered Bloom insert of
//                gen_block_hash (sdbf_core.cc:142-191) on one wave with the
//                filter spread one 32-bit word per lane.
//      stream mode: candidates (score > 16) are SHA-1 hashed in parallel and
//                written per tile; stream_insert_kernel then replays
//                gen_chunk_hash (sdbf_core.cc:94-136) in order, one wave per
//                input, with the 16 KiB de-duplication filter in LDS.
//
// Compare: one wave per (64 reference-filter instances, target chunk); each
// lane holds its reference filter in 64 VGPRs and the target filters stream
// through SGPRs (uniform scalar loads).  AND + v_bcnt with the early-exit
// checkpoints of bf_bitcount_cut_256 (sdbf/bf_utils.cc:115-206) and the score
// of sdbf_max_score (sdbf_core.cc:396-455), keeping the maximum as an exact
// fraction.  pair_sum_kernel reproduces sdbf_score's ordered double sum
// (sdbf_core.cc:331-389).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "sdbf_tables.h"
#include "sdbfx_internal.h"

#define RANK_LANES 128
#define RUN_CHUNKS 9     // 16-window chunks per ranks lane: 128 * 9 * 16 >= LMAX
#define LMAX (SDBFX_TILE + 64 + SDBFX_WARM + 64 + 8)
// ranks runs cover every held position, and their 16-position chunk writes
// stay inside P (a tile holds at most LMAX - 8 positions)
static_assert(((LMAX + RANK_LANES - 1) / RANK_LANES + 15) / 16 <= RUN_CHUNKS, "ranks run too long");
static_assert((LMAX - 8 + 15) / 16 * 16 <= LMAX, "ranks chunk writes past P");
// P stores position r (relative to the tile's base) at index pix(r): 8 bytes
// of padding after every 64 positions, so lanes that each walk their own
// 64-position block (F) or 64-step segment (scores) hit different banks.
// Groups of 4 positions starting at a multiple of 4 stay contiguous.
#define PIX(r) ((r) + (((r) >> 6) << 2))
#define PSIZE (LMAX + 4 * ((LMAX + 63) / 64))
__device__ __forceinline__ uint32_t pix(uint32_t r) { return PIX(r); }
#define BCAND_CAP 1152   // block mode: positions with score >= 15 in <= 16320 positions
#define NWAVE (SDBFX_NT / 64)

__constant__ uint16_t c_ranks[SDBFX_RANKS_LEN] = {SDBFX_ENTR64_RANKS_INIT};
__constant__ uint16_t c_cut[SDBFX_CUTOFFS_LEN] = {SDBFX_CUTOFFS256_INIT};
__constant__ uint16_t c_cut64[SDBFX_CUTOFFS_LEN] = {SDBFX_CUTOFFS64_INIT};
__constant__ uint32_t c_entropy[65] = {SDBFX_ENTROPY64_INIT};

struct __align__(16) TileSmem {
    uint16_t P[PSIZE];                      // rank | (F - i) << 10, positions [base, base+L) at pix()
    union {
        uint8_t hist[RANK_LANES * 256];     // ranks phase: one byte histogram per lane
        struct {
            uint32_t score[SDBFX_TILE / 4]; // owned positions' scores, 4 per word
            uint32_t cand[BCAND_CAP];       // candidates: position (relative to t0) | score << 16
            uint32_t hsha[5][SDBFX_NT];     // block mode: SHA-1 words of a batch
            uint32_t csc[SDBFX_NT];         // block mode: scores of a batch
        } s;
    } u;
    uint16_t rtab[SDBFX_NT * ((SDBFX_RANKS_LEN + SDBFX_NT - 1) / SDBFX_NT)];   // ENTR64_RANKS, tail padded
    int32_t dtab[SDBFX_NT];                 // E[c] - E[c-1] for c in 1..64; 0 elsewhere
    uint32_t start_m[SDBFX_NT];
    uint32_t end_m[SDBFX_NT];
    uint32_t histo[68];
    uint32_t filt[64];                      // block mode: the filter being built
    uint32_t chunk_cnt[64];
    uint32_t chunk_off[64];
    uint32_t misc[16];
};

// ------------------------------------------------------------------ SHA-1

__device__ __forceinline__ uint32_t rol32(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

// FIPS 180-4 SHA-1 compression of one 512-bit block.
__device__ __forceinline__ void sha1_compress(uint32_t H[5], uint32_t w[16])
{
    uint32_t a = H[0], b = H[1], c = H[2], d = H[3], e = H[4];
#pragma unroll
    for (int t = 0; t < 80; t++) {
        uint32_t wt;
        if (t < 16) {
            wt = w[t];
        } else {
            wt = rol32(w[(t - 3) & 15] ^ w[(t - 8) & 15] ^ w[(t - 14) & 15] ^ w[t & 15], 1);
            w[t & 15] = wt;
        }
        uint32_t f, k;
        if (t < 20) {
            f = (b & c) | (~b & d);
            k = 0x5A827999u;
        } else if (t < 40) {
            f = b ^ c ^ d;
            k = 0x6ED9EBA1u;
        } else if (t < 60) {
            f = (b & c) | (b & d) | (c & d);
            k = 0x8F1BBCDCu;
        } else {
            f = b ^ c ^ d;
            k = 0xCA62C1D6u;
        }
        uint32_t tmp = rol32(a, 5) + f + e + k + wt;
        e = d;
        d = c;
        c = rol32(b, 30);
        b = a;
        a = tmp;
    }
    H[0] += a;
    H[1] += b;
    H[2] += c;
    H[3] += d;
    H[4] += e;
}

// SHA-1 of the 64 bytes at p (any alignment; the 4 bytes after p+63 must be
// readable).  out[k] is digest word k read as a little-endian uint32, which is
// how the reference consumes the digest (sdbf_core.cc:105-106).
__device__ __forceinline__ void sha1_64(const uint8_t* p, uint32_t out[5])
{
    const uintptr_t ad = (uintptr_t)p;
    const uint32_t* q = (const uint32_t*)(ad & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(ad & 3);
    uint32_t d[17];
#pragma unroll
    for (int k = 0; k < 16; k++)
        d[k] = q[k];
    d[16] = sh ? q[16] : 0u;
    uint32_t w[16];
#pragma unroll
    for (int k = 0; k < 16; k++)
        w[k] = __builtin_bswap32(__builtin_amdgcn_alignbyte(d[k + 1], d[k], sh));
    uint32_t H[5] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u, 0xC3D2E1F0u};
    sha1_compress(H, w);
    uint32_t pad[16] = {0x80000000u, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 512u};
    sha1_compress(H, pad);
#pragma unroll
    for (int k = 0; k < 5; k++)
        out[k] = __builtin_bswap32(H[k]);
}

// ------------------------------------------------------------------ helpers

// Compiler controls; they change no result.  opaque() hides a value's origin
// (so that a prefetch is not merged with the load it feeds) and makes it
// needed at that point (so that a load issued early is not moved down to its
// use).  The CPU emulation of the kernels (tests/hipemu, which defines
// HIPEMU) uses the plain forms.
#if defined(HIPEMU)
__device__ __forceinline__ void opaque(uint32_t&) {}
__device__ __forceinline__ void opaque(uint64_t&) {}
#else
__device__ __forceinline__ void opaque(uint32_t& x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void opaque(uint64_t& x) { asm volatile("" : "+v"(x)); }
#endif
__device__ __forceinline__ void opaque(uint4& v)
{
    opaque(v.x);
    opaque(v.y);
    opaque(v.z);
    opaque(v.w);
}

__device__ __forceinline__ uint32_t rank_at(const uint16_t* P, uint32_t base, uint32_t pos)
{
    return P[pix(pos - base)] & 0x3FFu;
}

// Canonical state before step i: the remembered minimum m matters only while
// the cheap slide is possible (m > i and rank(m) > 0); every other state makes
// step i a full rescan and is written 0.
__device__ __forceinline__ uint32_t canon(const uint16_t* P, uint32_t base, uint32_t i, uint32_t m)
{
    return (m > i && rank_at(P, base, m) > 0) ? m : 0u;
}

__device__ __forceinline__ uint32_t score_at(const TileSmem& sm, uint32_t rel)
{
    return ((const uint8_t*)sm.u.s.score)[rel];
}

// Wave-wide scans with DPP row shifts (16-lane rows) and row read-lanes.
#define DPP_MIN(v, ctrl) v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFFu, (int)(v), ctrl, 0xF, 0xF, false))
#define DPP_ADD(v, ctrl) v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(v), ctrl, 0xF, 0xF, false)

// inclusive minimum over lanes [0, lane]
__device__ __forceinline__ uint32_t wave_prefix_min(uint32_t v, int lane)
{
    DPP_MIN(v, 0x111);
    DPP_MIN(v, 0x112);
    DPP_MIN(v, 0x114);
    DPP_MIN(v, 0x118);
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
    const uint32_t p2 = min(r0, r1), p3 = min(p2, r2);
    const int row = lane >> 4;
    const uint32_t add = row == 0 ? 0xFFFFFFFFu : row == 1 ? r0 : row == 2 ? p2 : p3;
    return min(v, add);
}

// inclusive minimum over lanes [lane, 63]
__device__ __forceinline__ uint32_t wave_suffix_min(uint32_t v, int lane)
{
    DPP_MIN(v, 0x101);
    DPP_MIN(v, 0x102);
    DPP_MIN(v, 0x104);
    DPP_MIN(v, 0x108);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    const uint32_t s1 = min(r2, r3), s0 = min(r1, s1);
    const int row = lane >> 4;
    const uint32_t add = row == 3 ? 0xFFFFFFFFu : row == 2 ? r3 : row == 1 ? s1 : s0;
    return min(v, add);
}

// exclusive sum over lanes [0, lane); *total gets the wave's sum
__device__ __forceinline__ uint32_t wave_excl_sum(uint32_t v, int lane, uint32_t* total)
{
    uint32_t x = v;
    DPP_ADD(x, 0x111);
    DPP_ADD(x, 0x112);
    DPP_ADD(x, 0x114);
    DPP_ADD(x, 0x118);
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)x, 15);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)x, 31);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)x, 47);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    const int row = lane >> 4;
    const uint32_t add = row == 0 ? 0u : row == 1 ? r0 : row == 2 ? r0 + r1 : r0 + r1 + r2;
    *total = r0 + r1 + r2 + r3;
    return x + add - v;
}

// number of set bits of m below this lane's bit
__device__ __forceinline__ uint32_t lanes_below(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// LDS written by some lanes of a wave, then read by others
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// position of the r-th (from 0) set bit of m; m has more than r set bits
__device__ __forceinline__ uint32_t select_bit(unsigned long long m, uint32_t r)
{
    uint32_t pos = 0, x = (uint32_t)m;
    uint32_t c = __popc(x);
    if (r >= c) {
        r -= c;
        x = (uint32_t)(m >> 32);
        pos = 32;
    }
#pragma unroll
    for (int w = 16; w >= 1; w >>= 1) {
        c = __popc(x & ((1u << w) - 1));
        if (r >= c) {
            r -= c;
            x >>= w;
            pos += w;
        }
    }
    return pos;
}

// Runs the score state machine over steps [i0, i1) from state m (i0 - base
// is a multiple of 4).  With CREDIT each step adds SIGN to the score of the
// position it elects, when that position is owned ([t0, t0+T)), with one LDS
// atomic per step (adding a run of equal positions at once needs a branch on
// the change of position, and some lane of the wave takes it at most steps).
// With HAND, the canonical state
// before step hand_i+1 is stored to sm.misc[2] when step hand_i runs.  Returns
// the state after step i1-1.  The words a step reads do not depend on the
// state, so they are loaded ahead of the transitions: the two words of each
// group of four steps two groups ahead, the four ranks at the groups' F
// positions one group ahead (group indices clamped to the last group).  Only
// the last group, when it is partial, checks step bounds.
template <int SIGN, bool CREDIT, bool HAND>
__device__ __forceinline__ uint32_t sm_run(TileSmem& sm, uint32_t base, uint32_t t0, uint32_t T, uint32_t i0,
                                           uint32_t i1, uint32_t m, uint32_t hand_i)
{
    if (i0 >= i1)
        return m;
    const uint16_t* P = sm.P;
    uint32_t rm = m ? rank_at(P, base, m) : 0u;
    const uint32_t ilast = i0 + ((i1 - i0 - 1) & ~3u);   // first step of the last group
    auto words = [&](uint32_t i, uint64_t& W, uint64_t& R) {
        const uint32_t x = pix(i - base);
        W = *(const uint64_t*)(P + x);
        R = *(const uint64_t*)(P + x + 68);   // pix(i + 64 - base)
    };
    auto franks = [&](uint32_t i, uint64_t W, uint32_t rf[4]) {
#pragma unroll
        for (int j = 0; j < 4; j++)
            rf[j] = P[pix(i + j + (((uint32_t)(W >> (16 * j)) & 0xFFFFu) >> 10) - base)];
    };
    // steps [i, i + nv) of one group
    auto group = [&](uint32_t i, uint64_t W4, uint64_t R4, const uint32_t rfw[4], uint32_t nv) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t ii = i + j;
            if ((uint32_t)j < nv) {
                // cheap slide (sdbf_core.cc:61-68) or full rescan (:69-81)
                const uint32_t w = (uint32_t)(W4 >> (16 * j)) & 0xFFFFu;
                const uint32_t r64 = (uint32_t)(R4 >> (16 * j)) & 0x3FFu;
                const uint32_t rf = rfw[j] & 0x3FFu;
                const bool cheap = ii > 0 && rm > 0 && r64 >= rm && ii < m;
                const uint32_t mc = r64 == rm ? ii + 64 : m;
                m = cheap ? mc : ii + (w >> 10);
                const bool has = cheap || rf != 0;
                rm = cheap ? rm : rf;
                if (CREDIT && has) {
                    const uint32_t rel = m - t0;
                    if (m >= t0 && rel < T)
                        atomicAdd(&sm.u.s.score[rel >> 2], ((uint32_t)SIGN) << (8 * (rel & 3)));
                }
                if (HAND && ii == hand_i)
                    sm.misc[2] = canon(P, base, ii + 1, m);
            }
        }
    };
    // Two groups per iteration with two sets of rank registers, so that no
    // register still being loaded is copied (a copy would wait for it).
    uint64_t W0, R0, W1, R1, W2, R2;
    uint32_t rfa[4], rfb[4];
    words(i0, W0, R0);
    franks(i0, W0, rfa);
    words(min(i0 + 4, ilast), W1, R1);
    uint32_t i = i0;
    for (; i + 4 < ilast; i += 8) {
        words(min(i + 8, ilast), W2, R2);
        franks(i + 4, W1, rfb);
        __builtin_amdgcn_sched_barrier(0);
        group(i, W0, R0, rfa, 4);
        opaque(W2);
        opaque(R2);
        words(min(i + 12, ilast), W0, R0);
        franks(i + 8, W2, rfa);
        __builtin_amdgcn_sched_barrier(0);
        group(i + 4, W1, R1, rfb, 4);
        opaque(W0);
        opaque(R0);
        W1 = W0;
        R1 = R0;
        W0 = W2;
        R0 = R2;
    }
    // one or two groups left: [i, i+4) and possibly the last group
    if (i < ilast) {
        franks(i + 4, W1, rfb);
        group(i, W0, R0, rfa, 4);
        i += 4;
        W0 = W1;
        R0 = R1;
#pragma unroll
        for (int j = 0; j < 4; j++)
            rfa[j] = rfb[j];
    }
    group(i, W0, R0, rfa, i1 - i);
    return m;
}

// Corrects a run of steps [i0, i1) that sm_run made from state mo when the
// true start state is mn: both trajectories advance together over the same
// words, and where a step's credits differ the old one is withdrawn and the
// new one added.  Once a lane's two trajectories reach the same canonical
// state their futures are identical (credits included), so the lane stops;
// the loop ends when every lane of the wave has stopped.  Returns the
// canonical state after step i1-1 of the new trajectory (the old run's end
// state when they merged).  With HAND, the canonical state before step
// hand_i+1 of the new trajectory is stored to sm.misc[2] if that step is
// re-run.  i0 - base is a multiple of 4.  Called by every lane of the wave;
// lanes with `run` false do nothing.
template <bool HAND>
__device__ __forceinline__ uint32_t sm_dual(TileSmem& sm, uint32_t base, uint32_t t0, uint32_t T, uint32_t i0,
                                            uint32_t i1, uint32_t mo, uint32_t mn, uint32_t old_end, uint32_t hand_i,
                                            bool run)
{
    const uint16_t* P = sm.P;
    // (the states of lanes that do not run may be unset)
    uint32_t ro = run && mo ? rank_at(P, base, mo) : 0u;
    uint32_t rn = run && mn ? rank_at(P, base, mn) : 0u;
    auto credit = [&](uint32_t pos, uint32_t v) {
        const uint32_t rel = pos - t0;
        if (pos >= t0 && rel < T)
            atomicAdd(&sm.u.s.score[rel >> 2], v << (8 * (rel & 3)));
    };
    bool live = run && i0 < i1, merged = false;
    uint32_t i = i0;
    while (__ballot(live)) {
        if (live) {
            const uint32_t x = pix(i - base);
            const uint64_t W = *(const uint64_t*)(P + x);
            const uint64_t R = *(const uint64_t*)(P + x + 68);   // pix(i + 64 - base)
            uint32_t rf[4];
#pragma unroll
            for (int j = 0; j < 4; j++)
                rf[j] = P[pix(i + j + (((uint32_t)(W >> (16 * j)) & 0xFFFFu) >> 10) - base)] & 0x3FFu;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t ii = i + j;
                if (ii < i1) {
                    const uint32_t w = (uint32_t)(W >> (16 * j)) & 0xFFFFu;
                    const uint32_t r64 = (uint32_t)(R >> (16 * j)) & 0x3FFu;
                    const uint32_t mf = ii + (w >> 10);
                    const bool co = ii > 0 && ro > 0 && r64 >= ro && ii < mo;
                    const bool cn = ii > 0 && rn > 0 && r64 >= rn && ii < mn;
                    mo = co ? (r64 == ro ? ii + 64 : mo) : mf;
                    mn = cn ? (r64 == rn ? ii + 64 : mn) : mf;
                    const bool ho = co || rf[j] != 0, hn = cn || rf[j] != 0;
                    ro = co ? ro : rf[j];
                    rn = cn ? rn : rf[j];
                    if (mo != mn || ho != hn) {
                        if (ho)
                            credit(mo, 0xFFFFFFFFu);   // -1 in the position's byte
                        if (hn)
                            credit(mn, 1u);
                    }
                    if (HAND && ii == hand_i)
                        sm.misc[2] = canon(P, base, ii + 1, mn);
                }
            }
            i += 4;
            const uint32_t ko = (mo > i && ro > 0) ? mo : 0u, kn = (mn > i && rn > 0) ? mn : 0u;
            if (ko == kn && i <= i1)
                merged = true;
            if (merged || i >= i1)
                live = false;
        }
    }
    return merged ? old_end : canon(P, base, i1, mn);
}

__device__ __forceinline__ uint32_t byte_of(const uint4& v, int j)
{
    const uint32_t w = (j < 4) ? v.x : (j < 8) ? v.y : (j < 12) ? v.z : v.w;
    return (w >> (8 * (j & 3))) & 0xFFu;
}

// ------------------------------------------------------------------ tile core

// Computes ranks, F and the scores of owned positions [t0, t0+T) of a unit of
// n bytes at u (n > 64, t0 + T <= n - 64, t0 == 0 or t0 >= 16384; u is 16-byte
// aligned and 16 bytes past n are readable).  forced >= 0 gives the exact
// canonical state before step t0-64.  When hand_i is not ~0, the canonical
// state before step hand_i+1 is left in sm.misc[2].  Returns 1 when every
// score is exact (certified), 0 when the incoming state could not be proven.
__device__ __forceinline__ uint32_t tile_core(TileSmem& sm, const uint8_t* __restrict__ u, uint32_t n,
                                              uint32_t t0, uint32_t T, int64_t forced, uint32_t hand_i)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const uint32_t a = t0 >= 64 ? t0 - 64 : 0;
    const uint32_t t_end = t0 + T;
    const uint32_t base = a >= SDBFX_WARM ? a - SDBFX_WARM : 0;
    const uint32_t L = t_end + 64 - base;   // positions held
    const uint32_t NI = t_end - base;       // steps whose F is needed
    const uint32_t lim = n - 64;            // positions with a full window

    // ---- 1. ranks: each lane rolls a byte histogram over a run of at most
    // RUN_CHUNKS * 16 windows that starts at a multiple of 16; the run's bytes
    // and the 64 after it are loaded into registers before the first use.
    // The count of byte value b for lane t is byte t & 3 of word
    // b * RANK_LANES / 4 + t / 4: the increment and the byte a lane reads
    // back are fixed per lane and the address is one multiply-add (a layout
    // with the byte chosen by b & 3 keeps every lane of a wave on its own
    // bank but needs four more instructions per update, and the ranks phase
    // is bound by instructions).  Lanes sharing a word change different
    // bytes, and no count passes 64, so the bytes never carry into each
    // other, whatever order the LDS applies one instruction's lanes in.
    // Counts change by word atomics that return the word
    // before the change: the LDS applies one wave's operations in order, so
    // the 32 updates of a 16-window chunk are all in flight at once and each
    // still sees every update before it.  Removing the outgoing byte before
    // adding the incoming one makes a step whose two bytes are equal change
    // the entropy by dtab[c] - dtab[c] = 0, as it must.  Past the last window
    // of the run the last chunk keeps sliding over real input bytes (at most
    // 79 bytes past it, within the 16 readable bytes past n); those steps only
    // touch the lane's own histogram and their ranks are replaced by 0.  The
    // loads (chunk indices past the run clamped to its last needed chunk) are
    // issued before the tables are filled, so their latency overlaps the fill.
    const uint32_t PR = ((L + RANK_LANES - 1) / RANK_LANES + 15) & ~15u;
    // The ranks lanes are waves 0-1 of even workgroups and waves 2-3 of odd
    // ones, so that two workgroups sharing a CU keep different SIMDs busy.
    const uint32_t rl = ((blockIdx.x >> 8) & 1) ? (uint32_t)tid - (SDBFX_NT - RANK_LANES) : (uint32_t)tid;
    const bool rank_lane = rl < RANK_LANES;
    const uint32_t r0 = rank_lane ? min(rl * PR, L) : L;
    const uint32_t r1 = min(r0 + PR, L);
    const uint32_t p0 = base + r0;
    const uint32_t pv = max(p0, min(base + r1, lim));   // windows exist for [p0, pv)
    const uint32_t nch = (pv - p0 + 15) / 16;
    // Every thread loads (no branches around the loads, so waiting for the
    // tables does not wait for the run); threads without a run read the
    // unit's first 64 bytes.
    constexpr int RT = (SDBFX_RANKS_LEN + SDBFX_NT - 1) / SDBFX_NT;
    uint32_t rt[RT];
#pragma unroll
    for (int k = 0; k < RT; k++)
        rt[k] = c_ranks[min(tid + k * SDBFX_NT, SDBFX_RANKS_LEN - 1)];
    const int32_t dt = (int32_t)c_entropy[min(tid, 64)] - (int32_t)c_entropy[min(max(tid, 1) - 1, 63)];
    const bool has_run = rank_lane && p0 < pv;
    const uint4* q0 = (const uint4*)(u + (has_run ? p0 : 0));
    const uint32_t lastc = has_run ? nch + 3 : 3;
    uint4 X[RUN_CHUNKS + 4];
#pragma unroll
    for (int c = 0; c < RUN_CHUNKS + 4; c++)
        X[c] = q0[min((uint32_t)c, lastc)];

#pragma unroll
    for (int k = 0; k < RT; k++)
        sm.rtab[tid + k * SDBFX_NT] = rt[k];
    sm.dtab[tid] = tid && tid < 65 ? dt : 0;
    for (int q = tid; q < RANK_LANES * 256 / 16; q += SDBFX_NT)
        ((uint4*)sm.u.hist)[q] = make_uint4(0, 0, 0, 0);
    __syncthreads();

    if (rank_lane) {
        uint32_t* H = (uint32_t*)sm.u.hist + (rl >> 2);
        const uint32_t hsh = 8 * (rl & 3), hinc = 1u << hsh;   // this lane's byte of a word
        if (p0 < pv) {
            // the first window's 64 increments (in two halves) and its
            // entropy from the counts they returned
            int32_t e = 0;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                uint32_t w[32];
#pragma unroll
                for (int k = 0; k < 32; k++)
                    w[k] = atomicAdd(H + byte_of(X[2 * h + (k >> 4)], k & 15) * (RANK_LANES / 4), hinc);
#pragma unroll
                for (int k = 0; k < 32; k++)
                    e += sm.dtab[((w[k] >> hsh) & 0xFFu) + 1];
            }
            auto update = [&](int c, uint32_t* wo_, uint32_t* wi_) {
#pragma unroll
                for (int j = (c == 0); j < 16; j++) {
                    const uint32_t bo = j ? byte_of(X[c], j - 1) : byte_of(X[c - 1], 15);
                    const uint32_t bi = j ? byte_of(X[c + 4], j - 1) : byte_of(X[c + 3], 15);
                    wo_[j] = atomicSub(H + bo * (RANK_LANES / 4), hinc);
                    wi_[j] = atomicAdd(H + bi * (RANK_LANES / 4), hinc);
                }
            };
            uint16_t* Pc = sm.P + pix(r0);
#pragma unroll
            for (int c = 0; c < RUN_CHUNKS; c++) {
                if ((uint32_t)c < nch) {
                    uint32_t wo_[16], wi_[16];
                    update(c, wo_, wi_);
                    uint32_t rk[8];
#pragma unroll
                    for (int j = 0; j < 16; j++) {
                        if (c > 0 || j > 0)
                            e += sm.dtab[((wi_[j] >> hsh) & 0xFFu) + 1] - sm.dtab[(wo_[j] >> hsh) & 0xFFu];
                        // windows from pv on get their rank 0 below
                        const uint32_t rv = sm.rtab[e >> 10];
                        rk[j >> 1] = (j & 1) ? (rk[j >> 1] | rv << 16) : rv;
                    }
                    uint16_t* pc = Pc + 16 * c + 4 * ((r0 + 16 * c) >> 6) - 4 * (r0 >> 6);   // pix(r0 + 16c)
#pragma unroll
                    for (int g = 0; g < 4; g++)
                        *(uint64_t*)(pc + 4 * g) = (uint64_t)rk[2 * g] | (uint64_t)rk[2 * g + 1] << 32;
                }
            }
        }
        // positions without a window: rank 0 (this also clears what the last
        // chunk wrote past pv, which is always inside the lane's own run)
        for (uint32_t p = max(pv, p0); p < base + r1; p++)
            sm.P[pix(p - base)] = 0;
    }
    __syncthreads();

    // ---- 2. F (and clear the score words, which alias the histograms).
    // One lane per 64-aligned block b.  The window starting at offset o of
    // block b takes the smaller rank of the suffix minimum of block b over
    // [o, 63] and the prefix minimum of block b+1 over [0, o-1] (the suffix on
    // ties: it is the earlier position); F is the last position of the run of
    // equal ranks that starts at that minimum, at most o + 63.  With da / db
    // the run-start bits of blocks b / b+1, the run ends one before the first
    // run start after the minimum.  A descending pass carries the suffix
    // minimum together with that distance (the next run start after each
    // offset is carried down the block); the ascending pass carries the prefix
    // minimum and its distance, found from db by shifts that are constants of
    // the unrolled loop.  Run-start bits past o + 64 never shorten a distance
    // below the cap of 63, so a sentinel bit at db's top keeps every search
    // bounded.
    for (int q = tid; q < SDBFX_TILE / 4; q += SDBFX_NT)
        sm.u.s.score[q] = 0;
    for (uint32_t blk = tid; blk * 64 < NI; blk += SDBFX_NT) {
        const uint32_t p0 = blk * 64;
        uint16_t* Pa = sm.P + pix(p0);
        const uint16_t* Pb = Pa + 68;   // pix(p0 + 64)
        uint64_t rw[16];
        uint64_t da = 0, db = 0;
        uint32_t prev = p0 ? (Pa[-5] & 0x3FFu) : 0xFFFFu;   // pix(p0 - 1)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            rw[q] = *(const uint64_t*)(Pa + 4 * q);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t r = (uint32_t)(rw[q] >> (16 * j)) & 0x3FFu;
                da |= (uint64_t)(r != prev) << (4 * q + j);
                prev = r;
            }
        }
        // block b+1: positions past L read as rank 0 and start no run
        const uint32_t nvb = L > p0 + 64 ? min(L - p0 - 64, 64u) : 0u;   // valid positions
        auto group_b = [&](int q) {
            const int32_t rem = (int32_t)nvb - 4 * q;
            const uint64_t keep = rem >= 4 ? ~0ull : rem <= 0 ? 0ull : (~0ull >> (64 - 16 * rem));
            return rem > 0 ? *(const uint64_t*)(Pb + 4 * q) & keep : 0ull;
        };
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const uint64_t vb = group_b(q);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t rb = (uint32_t)(vb >> (16 * j)) & 0x3FFu;
                db |= (uint64_t)(rb != prev) << (4 * q + j);
                prev = rb;
            }
        }
        db &= nvb >= 64 ? ~0ull : (1ull << nvb) - 1;
        // the passes below read bits and ranks back from these words rather
        // than keeping 64 unpacked ranks and comparisons live
        opaque(da);
        opaque(db);
#pragma unroll
        for (int q = 0; q < 16; q++)
            opaque(rw[q]);
        // Minima are taken over packed words rank' << 14 | offset << 8 | t,
        // rank' = (rank - 1) mod 1024 (rank 0, which never wins, becomes
        // 1023), t the first run start after that offset: one min selects
        // the minimum and carries its t, with no lane masks.
        // descending: suffix minimum rank' | distance from o to its run's end
        uint32_t suf2[32];
        {
            uint32_t sfx = 0xFFFFFFFFu;
            uint32_t nrs = db ? 64 + (uint32_t)__builtin_ctzll(db) : 128u;   // first run start after 63
#pragma unroll
            for (int o = 63; o >= 0; o--) {
                if (o < 63) {
                    const uint32_t bit = (uint32_t)(da >> (o + 1)) & 1u;
                    nrs = min(nrs, (uint32_t)(o + 1 + 256) - (bit << 8));
                }
                const uint32_t r = (uint32_t)(rw[o >> 2] >> (16 * (o & 3))) & 0x3FFu;
                sfx = min(sfx, (((r + 1023u) & 0x3FFu) << 14) | ((uint32_t)o << 8) | nrs);
                const uint32_t v = (sfx >> 14) | (min((sfx & 0xFFu) - 1 - (uint32_t)o, 63u) << 10);
                if (o & 1)
                    suf2[o >> 1] = v << 16;
                else
                    suf2[o >> 1] |= v;
            }
        }
        // ascending: the prefix minimum of block b+1 over [0, o-1]
        uint32_t pfx = 0xFFFFFFFFu;
        const uint64_t dbs = db | (1ull << 63);
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const uint64_t vb = group_b(q);   // read again: fewer live registers
            uint32_t out4[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int o = 4 * q + j;
                const uint32_t r = (uint32_t)(rw[q] >> (16 * j)) & 0x3FFu;
                const uint32_t sv = (o & 1) ? (suf2[o >> 1] >> 16) : (suf2[o >> 1] & 0xFFFFu);
                const uint32_t f =
                    (pfx >> 14) < (sv & 0x3FFu) ? min((pfx & 0xFFu) - 1 - (uint32_t)o, 63u) : (sv >> 10);
                out4[j] = r ? (r | (f << 10)) : 0u;
                if (o < 63) {
                    const uint32_t rb = (uint32_t)(vb >> (16 * j)) & 0x3FFu;
                    const uint32_t nb = 64 + o + 1 + (uint32_t)__builtin_ctzll(dbs >> (o + 1));
                    pfx = min(pfx, (((rb + 1023u) & 0x3FFu) << 14) | ((uint32_t)o << 8) | nb);
                }
            }
            if (p0 + 4 * q < NI) {
                const uint64_t v = (uint64_t)out4[0] | ((uint64_t)out4[1] << 16) | ((uint64_t)out4[2] << 32) |
                                   ((uint64_t)out4[3] << 48);
                *(uint64_t*)(Pa + 4 * q) = v;
            }
        }
    }
    __syncthreads();

    // ---- 3. scores
    const uint32_t NIter = t_end - a;
    const uint32_t S = ((NIter + SDBFX_NT - 1) / SDBFX_NT + 3) & ~3u;
    const uint32_t s = a + tid * S;
    const uint32_t e = min(s + S, t_end);
    const bool active = s < t_end;

    if (wave == 0) {
        uint32_t st = 0, cert = 1;
        if (forced >= 0) {
            st = (uint32_t)forced;
        } else if (a > 0) {
            // all 64 states possible before step `base` (a full rescan, or a
            // remembered minimum at base+1..base+63) run through the warm-up
            uint32_t m = lane ? base + lane : 0u;
            if (m && rank_at(sm.P, base, m) == 0)
                m = 0;
            m = sm_run<0, false, false>(sm, base, t0, T, base, a, m, 0);
            const uint32_t c = canon(sm.P, base, a, m);
            const uint32_t c0 = (uint32_t)__builtin_amdgcn_readlane((int)c, 0);
            cert = __ballot(c != c0) == 0 ? 1u : 0u;
            st = c0;
        }
        if (lane == 0) {
            sm.misc[0] = st;
            sm.misc[1] = cert;
        }
    }
    __syncthreads();

    uint32_t mstart = 0;
    if (tid == 0) {
        mstart = sm.misc[0];
    } else if (active) {
        const uint32_t w = (s > base + SDBFX_WARM) ? s - SDBFX_WARM : base;
        const uint32_t m = sm_run<0, false, false>(sm, base, t0, T, w, s, 0u, 0);
        mstart = canon(sm.P, base, s, m);
    }
    if (active) {
        const uint32_t m = hand_i != 0xFFFFFFFFu
                               ? sm_run<1, true, true>(sm, base, t0, T, s, e, mstart, hand_i)
                               : sm_run<1, true, false>(sm, base, t0, T, s, e, mstart, 0);
        sm.start_m[tid] = mstart;
        sm.end_m[tid] = canon(sm.P, base, e, m);
    }
    __syncthreads();
    for (;;) {
        bool mis = false;
        uint32_t pred = 0;
        if (tid > 0 && active) {
            pred = sm.end_m[tid - 1];
            mis = pred != sm.start_m[tid];
        }
        __syncthreads();
        {
            // every lane of the wave takes part in sm_dual's wave-wide loop;
            // the ones without a mismatch start out stopped
            const uint32_t old = sm.start_m[tid], old_end = sm.end_m[tid];
            const uint32_t ne = hand_i != 0xFFFFFFFFu
                                    ? sm_dual<true>(sm, base, t0, T, s, e, old, pred, old_end, hand_i, mis)
                                    : sm_dual<false>(sm, base, t0, T, s, e, old, pred, old_end, 0, mis);
            if (mis) {
                sm.end_m[tid] = ne;
                sm.start_m[tid] = pred;
            }
        }
        if (!__syncthreads_or(mis))
            break;
    }
    return sm.misc[1];
}

// Ordered compaction of the owned positions whose score is >= TH into
// sm.u.s.cand as (position | score << 16); returns the count.  Every score
// byte of the tile is read: bytes past the owned positions are 0 (the F
// phase clears them and credits stay inside), which TH >= 1 rejects.  Each
// wave takes 4096 positions as 4 chunks of 1024 in which lane l holds
// positions 16 l .. 16 l + 15 (one 16-byte LDS load), so that the order
// inside a chunk is lane order and one wave-wide scan per chunk places every
// lane.  A score byte b (<= 64) is >= TH exactly when bit 7 of
// b + 0x80 - TH is set, four bytes per 32-bit add.
template <uint32_t TH>
__device__ __forceinline__ uint32_t compact(TileSmem& sm)
{
    static_assert(TH >= 1 && TH <= 64, "threshold");
    constexpr uint32_t ADD = (0x80u - TH) * 0x01010101u;
    constexpr int NCH = SDBFX_TILE / 1024 / NWAVE;
    static_assert(NCH * 1024 * NWAVE == SDBFX_TILE, "chunks cover the tile");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint4 sw[NCH];
#pragma unroll
    for (int k = 0; k < NCH; k++)
        sw[k] = ((const uint4*)sm.u.s.score)[(wave * NCH + k) * 64 + lane];   // words 4 l .. 4 l + 3 of the chunk
    // per chunk: the lane's 16 flags (bit 4j + i: byte i of word j, the
    // position order), its place among the chunk's kept positions and the
    // chunk's count
    uint32_t fl[NCH], pre[NCH], ctot[NCH], mine = 0;
#pragma unroll
    for (int k = 0; k < NCH; k++) {
        const uint32_t w[4] = {sw[k].x, sw[k].y, sw[k].z, sw[k].w};
        uint32_t f = 0, c = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t t = (w[j] + ADD) & 0x80808080u;
            c = (uint32_t)__popc(t) + c;
            // bits 7, 15, 23, 31 to bits 21..24 (distinct partial products)
            f |= ((((t >> 7) * 0x00204081u) >> 21) & 0xFu) << (4 * j);
        }
        fl[k] = f;
        pre[k] = wave_excl_sum(c, lane, &ctot[k]);
        mine += ctot[k];
    }
    if (lane == 0)
        sm.chunk_cnt[wave] = mine;
    __syncthreads();
    uint32_t o = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; w++) {
        const uint32_t t = sm.chunk_cnt[w];
        o += w < wave ? t : 0u;
        total += t;
    }
#pragma unroll
    for (int k = 0; k < NCH; k++) {
        const uint32_t q0 = (uint32_t)(wave * NCH + k) * 1024 + (uint32_t)lane * 16;
        uint32_t at = o + pre[k], f = fl[k];
        while (f) {
            const uint32_t bit = (uint32_t)__builtin_ctz(f);
            const uint32_t j = bit >> 2;
            const uint32_t wv = j == 0 ? sw[k].x : j == 1 ? sw[k].y : j == 2 ? sw[k].z : sw[k].w;
            sm.u.s.cand[at++] = (q0 + bit) | (((wv >> (8 * (bit & 3))) & 0xFFu) << 16);
            f &= f - 1;
        }
        o += ctot[k];
    }
    __syncthreads();
    return total;
}

// ------------------------------------------------------------------ block mode

// Keeps the entries of sm.u.s.cand[0, n) whose score is >= k (k > 0), in
// order, at the front of the list; returns their number.  Each thread takes
// BCAND_CAP / SDBFX_NT consecutive entries (a stride coprime with the 64
// banks); all are read before any is written.
__device__ __forceinline__ uint32_t keep_cand(TileSmem& sm, uint32_t n, uint32_t k)
{
    constexpr int PER = (BCAND_CAP + SDBFX_NT - 1) / SDBFX_NT;
    static_assert(PER % 2 == 1, "entries per thread must be odd (bank stride)");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t v[PER], cnt = 0;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint32_t i = (uint32_t)tid * PER + j;
        v[j] = i < n ? sm.u.s.cand[i] : 0u;   // score 0 past the list: not kept
        cnt += (v[j] >> 16) >= k ? 1u : 0u;
    }
    uint32_t tot;
    uint32_t o = wave_excl_sum(cnt, lane, &tot);
    if (lane == 0)
        sm.chunk_cnt[wave] = tot;
    __syncthreads();
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; w++) {
        const uint32_t t = sm.chunk_cnt[w];
        o += w < wave ? t : 0u;
        total += t;
    }
#pragma unroll
    for (int j = 0; j < PER; j++)
        if ((v[j] >> 16) >= k)
            sm.u.s.cand[o++] = v[j];
    __syncthreads();
    return total;
}

// One workgroup per dd-block (sdbf_core.cc:252-282 per block, :305-319 tail).
// With feat (not null), the SHA-1 words of the block's counted features, the
// ones gen_block_hash hands to an index (sdbf_core.cc:163-177), go to
// feat[((blockIdx.x * 192) + k) * 5 + j] for the k-th counted feature, in
// order.
__global__ __launch_bounds__(SDBFX_NT, 2) void block_digest_kernel(
    const uint8_t* __restrict__ data, const BlockDesc* __restrict__ blocks, uint8_t* __restrict__ filters,
    uint16_t* __restrict__ elem_counts, uint16_t* __restrict__ hamming, uint32_t* __restrict__ feat)
{
    __shared__ TileSmem sm;
    const BlockDesc bd = blocks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint8_t* u = data + bd.off;
    const uint32_t n = bd.n;
    const uint32_t T = n - 64;

    tile_core(sm, u, n, 0, T, -1, 0xFFFFFFFFu);

    // The positions of [0, n-64) with score >= 15, the lowest threshold the
    // walk below can settle on, in order; then the histogram of their scores
    // >= 16 (the walk's counts) from that list.
    if (tid < 68)
        sm.histo[tid] = 0;
    const uint32_t C15 = compact<SDBFX_THRESHOLD - 1>(sm);
    for (uint32_t i = tid; i < C15; i += SDBFX_NT) {
        const uint32_t sc = sm.u.s.cand[i] >> 16;
        if (sc >= SDBFX_THRESHOLD)
            atomicAdd(&sm.histo[sc], 1u);
    }
    __syncthreads();
    if (wave == 0) {
        // the reference's downward walk k = 65, 64, ... (sdbf_core.cc:271-276)
        // stops at the first k whose count takes the running sum past the
        // maximum; lane l holds k = 65 - l and the sum of the counts above it
        uint32_t k = SDBFX_THRESHOLD, sum = 0;
        int32_t allowed = SDBFX_MAX_ELEM_DD;
        if (!bd.tail) {
            const bool in = lane <= 65 - SDBFX_THRESHOLD;
            const uint32_t h = in ? sm.histo[65 - lane] : 0u;
            uint32_t tot;
            const uint32_t above = wave_excl_sum(h, lane, &tot);
            const unsigned long long stop =
                __ballot(in && above <= SDBFX_MAX_ELEM_DD && above + h > SDBFX_MAX_ELEM_DD);
            if (stop) {
                const int first = __builtin_ctzll(stop);
                k = 65 - first;
                sum = (uint32_t)__builtin_amdgcn_readlane((int)above, first);
            } else {
                k = SDBFX_THRESHOLD - 1;
                sum = tot;
            }
            allowed = (int32_t)(SDBFX_MAX_ELEM_DD - sum);
        }
        if (lane == 0) {
            sm.misc[3] = k;
            sm.misc[4] = (uint32_t)allowed;
        }
    }
    __syncthreads();
    const uint32_t k = sm.misc[3];
    int32_t allowed = (int32_t)sm.misc[4];
    const uint32_t C = keep_cand(sm, C15, k);

    // Ordered insertion (gen_block_hash, sdbf_core.cc:157-183), 64 candidates
    // per step on wave 0.  Each bit of a batch is owned by the first selected
    // lane that sets it (LDS minimum); a lane counts when it owns a bit that
    // was clear.  Lanes after the first one whose selection was speculative
    // and wrong (allowance used up) or after the 192nd element are re-run.
    uint32_t* own = sm.u.s.score;   // 2048 entries; the score words are dead now
    if (tid < 64)
        sm.filt[tid] = 0;
    uint32_t hash_cnt = 0, next = 0;
    bool done = false;
    __syncthreads();
    while (!done && next < C) {
        const uint32_t wcount = min((uint32_t)SDBFX_NT, C - next);
        if ((uint32_t)tid < wcount) {
            const uint32_t ce = sm.u.s.cand[next + tid];
            uint32_t h[5];
            sha1_64(u + (ce & 0xFFFFu), h);
#pragma unroll
            for (int j = 0; j < 5; j++)
                sm.u.s.hsha[j][tid] = h[j];
            sm.u.s.csc[tid] = ce >> 16;
        }
        __syncthreads();
        if (wave == 0) {
            uint32_t wpos = 0;
            while (wpos < wcount && !done) {
                const uint32_t li = wpos + lane;
                const bool act = li < wcount;
                const uint32_t sc = act ? sm.u.s.csc[li] : 0u;
                uint32_t b[5];
#pragma unroll
                for (int j = 0; j < 5; j++)
                    b[j] = act ? (sm.u.s.hsha[j][li] & 0x7FFu) : 0u;
                const bool sel = act && (sc > k || (sc == k && allowed > 0));
                if (sel) {
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        own[b[j]] = 64;
                }
                __builtin_amdgcn_wave_barrier();
                if (sel) {
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        atomicMin(&own[b[j]], (uint32_t)lane);
                }
                __builtin_amdgcn_wave_barrier();
                bool nw = false;
                if (sel) {
#pragma unroll
                    for (int j = 0; j < 5; j++) {
                        const uint32_t fw = sm.filt[b[j] >> 5];
                        if (own[b[j]] == (uint32_t)lane && !((fw >> (b[j] & 31)) & 1u))
                            nw = true;
                    }
                }
                const uint64_t mc = __ballot(nw);
                const uint64_t mk = __ballot(nw && sc == k);
                const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
                const uint32_t before = (uint32_t)__popcll(mc & lt);
                const uint32_t kbefore = (uint32_t)__popcll(mk & lt);
                const uint64_t mv = __ballot(sel && sc == k && allowed - (int32_t)kbefore <= 0);
                const uint64_t mcap = __ballot(nw && hash_cnt + before + 1 == SDBFX_MAX_ELEM_DD);
                uint32_t r = min(64u, wcount - wpos);
                if (mv)
                    r = min(r, (uint32_t)__builtin_ctzll(mv));
                bool hit = false;
                if (mcap) {
                    const uint32_t r2 = (uint32_t)__builtin_ctzll(mcap) + 1;
                    if (r2 <= r) {
                        r = r2;
                        hit = true;
                    }
                }
                // commit lanes [0, r): their bits into the filter
                __builtin_amdgcn_wave_barrier();
                if (sel && (uint32_t)lane < r) {
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        atomicOr(&sm.filt[b[j] >> 5], 1u << (b[j] & 31));
                }
                if (feat && nw && (uint32_t)lane < r) {
                    uint32_t* fo = feat + ((uint64_t)blockIdx.x * SDBFX_MAX_ELEM_DD + hash_cnt + before) * 5;
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        fo[j] = sm.u.s.hsha[j][li];
                }
                __builtin_amdgcn_wave_barrier();
                const uint64_t mr = r >= 64 ? ~0ull : ((1ull << r) - 1);
                hash_cnt += (uint32_t)__popcll(mc & mr);
                allowed -= (int32_t)__popcll(mk & mr);
                wpos += r;
                done = hit;
            }
            if (lane == 0) {
                sm.misc[5] = done ? 1u : 0u;
                sm.misc[7] = next + wpos;
            }
        }
        __syncthreads();
        done = sm.misc[5] != 0;
        next = sm.misc[7];
        __syncthreads();
    }
    if (wave == 0) {
        const uint32_t fword = sm.filt[lane];
        ((uint32_t*)(filters + bd.out * SDBFX_BF))[lane] = fword;
        uint32_t pc = __popc(fword);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            pc += __shfl_down(pc, off);
        if (lane == 0) {
            elem_counts[bd.out] = (uint16_t)hash_cnt;
            hamming[bd.out] = (uint16_t)pc;
        }
    }
}

// The dd-blocks of one input of `len` bytes at offset 0 (sdbf_core.cc:
// 288-324): full blocks, then a tail of at least 512 bytes; block t is
// filter t.  One thread per block.
__global__ void block_desc_kernel(BlockDesc* __restrict__ out, uint64_t len, uint32_t dd, uint32_t nblk)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblk)
        return;
    const uint64_t qt = len / dd;
    BlockDesc b;
    if (t < qt) {
        b.off = (uint64_t)t * dd;
        b.n = dd;
        b.tail = 0;
    } else {
        b.off = qt * dd;
        b.n = (uint32_t)(len % dd);
        b.tail = 1;
    }
    b.out = t;
    out[t] = b;
}

// ------------------------------------------------------------------ stream mode

__device__ __forceinline__ void stream_tile(TileSmem& sm, const uint8_t* __restrict__ data, const TileDesc& td,
                            uint32_t tile, int64_t forced, uint32_t* __restrict__ cand,
                            TileState* __restrict__ tstate)
{
    const int tid = threadIdx.x;
    const uint8_t* u = data + td.unit_off;
    const uint32_t t_end = td.t0 + td.T;
    const uint32_t hand_i = (t_end < td.n - 64) ? t_end - 65 : 0xFFFFFFFFu;
    if (tid == 0)
        sm.misc[2] = 0;
    const uint32_t cert = tile_core(sm, u, td.n, td.t0, td.T, forced, hand_i);
    const uint32_t C = compact<SDBFX_THRESHOLD + 1>(sm);
    uint32_t* cw = cand + (uint64_t)tile * 5 * SDBFX_CAND_CAP;
    for (uint32_t c = tid; c < C; c += SDBFX_NT) {
        uint32_t h[5];
        sha1_64(u + td.t0 + (sm.u.s.cand[c] & 0xFFFFu), h);
#pragma unroll
        for (int j = 0; j < 5; j++)
            cw[j * SDBFX_CAND_CAP + c] = h[j];
    }
    if (tid == 0) {
        TileState st;
        st.count = C;
        st.handoff = sm.misc[2];
        st.certified = cert;
        st.pad = 0;
        tstate[tile] = st;
    }
    __syncthreads();
}

__global__ __launch_bounds__(SDBFX_NT, 2) void stream_tile_kernel(
    const uint8_t* __restrict__ data, const TileDesc* __restrict__ tiles, uint32_t* __restrict__ cand,
    TileState* __restrict__ tstate)
{
    __shared__ TileSmem sm;
    const TileDesc td = tiles[blockIdx.x];
    stream_tile(sm, data, td, blockIdx.x, -1, cand, tstate);
}

// A tile whose incoming state was not proven is re-run from the exact
// hand-off state of the tile before it (stream_tile with that state marks it
// certified).  Three passes: stream_fixup_plan_kernel marks the tiles whose
// predecessor is certified; stream_fixup_par_kernel re-runs those, all at
// once (their predecessors are not re-run, so their hand-offs do not change
// meanwhile); stream_fixup_kernel then re-runs what is left -- tiles whose
// predecessor was also unproven -- in order, one workgroup per unit.
__global__ void stream_fixup_plan_kernel(const TileDesc* __restrict__ tiles, uint32_t ntiles,
                                         TileState* __restrict__ tstate)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles)
        return;
    tstate[t].pad = (!tiles[t].first && !tstate[t].certified && tstate[t - 1].certified) ? 1u : 0u;
}

__global__ __launch_bounds__(SDBFX_NT) void stream_fixup_par_kernel(
    const uint8_t* __restrict__ data, const TileDesc* __restrict__ tiles, uint32_t* __restrict__ cand,
    TileState* __restrict__ tstate)
{
    __shared__ TileSmem sm;
    const uint32_t t = blockIdx.x;
    if (!tstate[t].pad)
        return;
    const uint32_t h = tstate[t - 1].handoff;
    stream_tile(sm, data, tiles[t], t, (int64_t)h, cand, tstate);
}

__global__ __launch_bounds__(SDBFX_NT) void stream_fixup_kernel(
    const uint8_t* __restrict__ data, const TileDesc* __restrict__ tiles, const UnitDesc* __restrict__ units,
    uint32_t* __restrict__ cand, TileState* __restrict__ tstate)
{
    __shared__ TileSmem sm;
    __shared__ uint32_t list[SDBFX_NT];   // unproven tiles of a run of SDBFX_NT, in order
    __shared__ uint32_t wcount[NWAVE];
    __shared__ uint32_t hand;
    const UnitDesc ud = units[blockIdx.x];
    const int wave = threadIdx.x >> 6;
    // the first tile of a unit starts at step 0 and is always exact
    for (uint32_t b0 = ud.tile_begin + 1; b0 < ud.tile_end; b0 += SDBFX_NT) {
        const uint32_t t = b0 + threadIdx.x;
        const bool unc = t < ud.tile_end && !((volatile TileState*)tstate)[t].certified;
        const unsigned long long m = __ballot(unc);
        if ((threadIdx.x & 63) == 0)
            wcount[wave] = (uint32_t)__popc((uint32_t)m) + (uint32_t)__popc((uint32_t)(m >> 32));
        __syncthreads();
        uint32_t pos = 0, total = 0;
        for (int w = 0; w < NWAVE; w++) {
            pos += w < wave ? wcount[w] : 0u;
            total += wcount[w];
        }
        if (unc)
            list[pos + lanes_below(m)] = t;
        __syncthreads();
        for (uint32_t k = 0; k < total; k++) {
            const uint32_t tt = list[k];
            // the predecessor is exact: certified from the start, re-run by
            // the parallel pass, or re-run just before in this loop
            if (threadIdx.x == 0)
                hand = ((volatile TileState*)tstate)[tt - 1].handoff;
            __syncthreads();
            const uint32_t h = hand;
            __syncthreads();
            stream_tile(sm, data, tiles[tt], tt, (int64_t)h, cand, tstate);
            __threadfence();
            __syncthreads();
        }
    }
}

// Ordered insertion of one stream-mode input (gen_chunk_hash per unit,
// sdbf_core.cc:94-136, with the de-duplication filter of sdbf_class.cc:378-380
// and bloom_filter.cc:334-359), then the final trim (sdbf_core.cc:235-238).
// One wave per input; the current 256-byte filter and the 16 KiB
// de-duplication filter are in LDS.
//
// Candidates are taken 64 at a time in order (lane l = the l-th).  Every one
// sets its 5 bits in the current filter (bf_sha1_insert); its count of new
// bits is the number of its distinct positions that were clear before the
// batch and that no lower lane of the batch also sets (first-lane ownership
// through atomicMin).  Those with new bits reach the de-duplication filter,
// where one is rejected when all its positions are set.  A position clear at
// the start of the batch and set by another lane of the batch is detected
// with two scratch bitmaps; when such positions could decide a rejection
// (all of a lane's clear positions are shared), the batch is replayed one
// candidate at a time.  The accepted candidates are counted in order; the
// batch is committed up to the one that completes a filter (160 elements)
// or a de-duplication filter (8738), and the rest is taken again with the
// new state.
//
// INDEXED (info->index set, the --index hashing of sdhash_threads.cc:376-431):
// between the filter and the de-duplication filter, every candidate with new
// filter bits is inserted into the index (bloom_filter::insert_sha1,
// bloom_filter.cc:312-359, in global memory at ixw with bit mask imask) and
// skipped when all its index bits were already set (sdbf_core.cc:110-115).
// Inputs share the index, so one wave takes all `nfiles` inputs in order.  A
// lane's index positions are read once per batch (loads that bypass the
// wave's L1, which the index atomics do not update); a position clear before
// the batch that another lane of the batch also sets is found with an LDS
// hash table of the batch's clear positions, and a lane whose clear positions
// are all shared sends the batch to the one-at-a-time replay, which inserts
// with atomics.  The index's element count (counted inserts) is carried in a
// register and written back to *ielem at the end.
struct InsertSmem {
    uint32_t big[SDBFX_BIG_BYTES / 4];      // de-duplication filter
    uint32_t seen[SDBFX_BIG_BYTES / 4];     // positions of the batch
    uint32_t twice[SDBFX_BIG_BYTES / 4];    // positions of the batch set by two lanes
    uint32_t own[SDBFX_BF * 8];             // lowest lane setting each filter bit
    uint32_t fw[SDBFX_BFW];                 // current filter
    uint32_t tinfo[64];                     // tiles [tw0, tw0 + 64): count | first << 31
};
#define IX_SLOTS 1024                        // index positions of one batch: at most 64 * 5
struct InsertSmemIx : InsertSmem {
    uint32_t ikey[IX_SLOTS];                // open addressing: an index position, or ~0
    uint32_t itw[IX_SLOTS];                 // 1: the position is set by two lanes
};

// index word w, read at the device's coherence point (not from the L1)
__device__ __forceinline__ uint32_t index_word(const uint32_t* ixw, uint32_t w)
{
    return __hip_atomic_load(ixw + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool INDEXED>
__device__ __forceinline__ void stream_insert_body(
    const StreamFile* __restrict__ files, uint32_t fbegin, uint32_t fend, const TileDesc* __restrict__ tiles,
    const TileState* __restrict__ tstate, const uint32_t* __restrict__ cand, uint8_t* __restrict__ stage,
    uint16_t* __restrict__ stage_ham, uint32_t* __restrict__ out_counts, uint32_t* __restrict__ ixw,
    uint32_t imask, unsigned long long* __restrict__ ielem)
{
    __shared__ typename std::conditional<INDEXED, InsertSmemIx, InsertSmem>::type ism;
    uint32_t* const big = ism.big;
    uint32_t* const fw = ism.fw;
    const int lane = threadIdx.x;
    unsigned long long ecount = 0;   // index elements (INDEXED)
    if constexpr (INDEXED) {
        ecount = *ielem;
        for (int q = lane; q < IX_SLOTS; q += 64) {
            ism.ikey[q] = 0xFFFFFFFFu;
            ism.itw[q] = 0;
        }
    }
    for (uint32_t fi = fbegin; fi < fend; fi++) {
    const StreamFile sf = files[fi];
    for (int q = lane; q < SDBFX_BIG_BYTES / 4; q += 64) {
        big[q] = 0;
        ism.seen[q] = 0;
        ism.twice[q] = 0;
    }
    for (int q = lane; q < SDBFX_BF * 8; q += 64)
        ism.own[q] = 0xFFFFFFFFu;
    fw[lane] = 0;
    __syncthreads();
    uint32_t bfc = 1, last = 0, bigfi = 0, overflow = 0;

    auto flush = [&](uint32_t idx) {
        const uint32_t word = fw[lane];
        uint32_t pc = __popc(word);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            pc += __shfl_down(pc, off);
        if (idx < sf.filt_cap) {
            ((uint32_t*)(stage + (sf.filt_base + idx) * SDBFX_BF))[lane] = word;
            if (lane == 0)
                stage_ham[sf.filt_base + idx] = (uint16_t)pc;
        } else {
            overflow = 1;
        }
    };
    // after an accepted candidate: rollover of the filter and of the
    // de-duplication filter (sdbf_core.cc:121-132)
    auto counted = [&](uint32_t k) {
        last += k;
        bigfi += k;
        if (last == SDBFX_MAX_ELEM) {
            flush(bfc - 1);
            wave_lds_sync();
            fw[lane] = 0;
            bfc++;
            last = 0;
        }
        if (bigfi == SDBFX_BIG_ELEM) {
            for (int q = lane; q < SDBFX_BIG_BYTES / 4; q += 64)
                big[q] = 0;
            bigfi = 0;
        }
        wave_lds_sync();
    };

    // The batches of 64 candidates in order: (tile t, first candidate c0).
    // Each batch's SHA-1 words are loaded one batch ahead, so that their
    // latency overlaps the batch before; a tile's restart flag (the first
    // tile of a 32 MiB unit resets the de-duplication count) is applied
    // when the first batch at or after it is taken.  The tiles' candidate
    // counts and restart flags are read 64 tiles at a time (one load per
    // lane) into LDS, so that moving to the next tile waits for no load.
    uint32_t tw0 = 0xFFFFFFFFu;
    auto tile_info = [&](uint32_t tt) {
        if (tw0 == 0xFFFFFFFFu || tt < tw0 || tt >= tw0 + 64) {
            const uint32_t k = tt + (uint32_t)lane;
            uint32_t v = 0;
            if (k < sf.tile_end)
                v = tstate[k].count | (tiles[k].first ? 0x80000000u : 0u);
            wave_lds_sync();
            ism.tinfo[lane] = v;
            wave_lds_sync();
            tw0 = tt;
        }
        return ism.tinfo[tt - tw0];
    };
    auto load_batch = [&](uint32_t tt, uint32_t cc, uint32_t CC, uint32_t h[5]) {
        const uint32_t* cw = cand + (uint64_t)tt * 5 * SDBFX_CAND_CAP;
#pragma unroll
        for (int j = 0; j < 5; j++)
            h[j] = (cc + lane < CC) ? cw[j * SDBFX_CAND_CAP + cc + lane] : 0u;
    };
    // the first tile from tt on with candidates (sf.tile_end if none), and its count
    auto next_tile = [&](uint32_t tt, uint32_t& CC) {
        CC = 0;
        while (tt < sf.tile_end && (CC = tile_info(tt) & 0x7FFFFFFFu) == 0)
            tt++;
        return tt;
    };
    uint32_t C;
    uint32_t t = next_tile(sf.tile_begin, C), c0 = 0, applied = sf.tile_begin;
    uint32_t hnext[5] = {0, 0, 0, 0, 0};
    if (t < sf.tile_end)
        load_batch(t, 0, C, hnext);
    while (t < sf.tile_end) {
        for (; applied <= t; applied++)
            if (tile_info(applied) >> 31)
                bigfi = 0;
        {
            const uint32_t cnt = min(64u, C - c0);
            uint32_t myh[5];
#pragma unroll
            for (int j = 0; j < 5; j++)
                myh[j] = hnext[j];
            // the next batch, loaded now
            uint32_t tn = t, c0n = c0 + 64, Cn = C;
            if (c0n >= C) {
                tn = next_tile(t + 1, Cn);
                c0n = 0;
            }
            if (tn < sf.tile_end)
                load_batch(tn, c0n, Cn, hnext);
            // positions and which of them repeat an earlier one of the lane
            uint32_t p[5], q[5], ip[5], pdup = 0, qdup = 0, idup = 0;
#pragma unroll
            for (int j = 0; j < 5; j++) {
                p[j] = myh[j] & 0x7FFu;
                q[j] = myh[j] & 0x1FFFFu;
                ip[j] = myh[j] & imask;
#pragma unroll
                for (int k = 0; k < j; k++) {
                    pdup |= (p[k] == p[j] ? 1u : 0u) << j;
                    qdup |= (q[k] == q[j] ? 1u : 0u) << j;
                    idup |= (ip[k] == ip[j] ? 1u : 0u) << j;
                }
            }
            uint32_t s = 0;   // first candidate of the chunk not yet committed
            while (s < cnt) {
                const bool act = (uint32_t)lane >= s && (uint32_t)lane < cnt;
                // 1. new bits in the current filter (a repeated position
                // takes a minimum with ~0, which changes nothing)
                if (act) {
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        atomicMin(&ism.own[p[j]], (pdup >> j) & 1 ? 0xFFFFFFFFu : (uint32_t)lane);
                }
                wave_lds_sync();
                // (the filter words and owners of all five positions are read
                // together, so that their LDS latencies overlap)
                uint32_t nnew = 0;
                if (act) {
                    uint32_t fv[5], ov[5];
#pragma unroll
                    for (int j = 0; j < 5; j++) {
                        fv[j] = fw[p[j] >> 5];
                        ov[j] = ism.own[p[j]];
                    }
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        nnew += (!((pdup >> j) & 1) && !((fv[j] >> (p[j] & 31)) & 1u) && ov[j] == (uint32_t)lane)
                                    ? 1u : 0u;
                }
                const bool reach = act && nnew > 0;
                // 1b. the index gate for those with new bits: positions clear
                // before the batch (iclr), and whether another lane of the
                // batch sets them too
                uint32_t iclr = 0;
                bool iamb = false;
                if constexpr (INDEXED) {
                    uint32_t slot[5];
                    if (reach) {
                        uint32_t iw[5];
#pragma unroll
                        for (int j = 0; j < 5; j++)
                            iw[j] = (idup >> j) & 1 ? 0xFFFFFFFFu : index_word(ixw, ip[j] >> 5);
#pragma unroll
                        for (int j = 0; j < 5; j++) {
                            slot[j] = 0;
                            if (!((iw[j] >> (ip[j] & 31)) & 1u)) {
                                iclr |= 1u << j;
                                uint32_t h = (ip[j] * 2654435761u) >> 22;   // 10 bits
                                for (;;) {
                                    const uint32_t o = atomicCAS(&ism.ikey[h], 0xFFFFFFFFu, ip[j]);
                                    if (o == 0xFFFFFFFFu)
                                        break;
                                    if (o == ip[j]) {
                                        atomicOr(&ism.itw[h], 1u);
                                        break;
                                    }
                                    h = (h + 1) & (IX_SLOTS - 1);
                                }
                                slot[j] = h;
                            }
                        }
                    }
                    wave_lds_sync();
                    uint32_t ishared = 0;
                    if (reach) {
#pragma unroll
                        for (int j = 0; j < 5; j++)
                            if (((iclr >> j) & 1) && ism.itw[slot[j]])
                                ishared |= 1u << j;
                    }
                    iamb = reach && iclr != 0 && ishared == iclr;
                    wave_lds_sync();
                    if (reach) {
#pragma unroll
                        for (int j = 0; j < 5; j++)
                            if ((iclr >> j) & 1) {
                                ism.ikey[slot[j]] = 0xFFFFFFFFu;
                                ism.itw[slot[j]] = 0;
                            }
                    }
                }
                // lanes that go on to the de-duplication filter
                const bool reach2 = INDEXED ? (reach && iclr != 0) : reach;
                // 2. the de-duplication filter for those
                // Every distinct position of the lane is marked in `seen` (and
                // in `twice` when another lane marked it first), clear or not:
                // whether a position was clear before the batch is the same
                // for every lane, so marks on set positions never make a
                // clear one look shared.  The five reads and the five marks
                // are issued together (an OR with 0 for a repeated position).
                uint32_t clr = 0;   // positions clear at the start of the batch
                uint32_t mk[5];     // the bit each position marks (0: repeated)
                if (reach2) {
                    uint32_t bv[5], old[5];
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        bv[j] = big[q[j] >> 5];
#pragma unroll
                    for (int j = 0; j < 5; j++) {
                        mk[j] = (qdup >> j) & 1 ? 0u : 1u << (q[j] & 31);
                        if (mk[j] && !(bv[j] & mk[j]))
                            clr |= 1u << j;
                        old[j] = atomicOr(&ism.seen[q[j] >> 5], mk[j]);
                    }
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        atomicOr(&ism.twice[q[j] >> 5], old[j] & mk[j]);
                }
                wave_lds_sync();
                uint32_t shared = 0;
                if (reach2) {
                    uint32_t tv[5];
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        tv[j] = ism.twice[q[j] >> 5];
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        if (((clr >> j) & 1) && ((tv[j] >> (q[j] & 31)) & 1u))
                            shared |= 1u << j;
                }
                // a rejection decided inside the batch: replay it in order
                const bool replay =
                    SDBFX_INSERT_REPLAY || __ballot((reach2 && clr != 0 && shared == clr) || iamb) != 0;
                const bool acc = reach2 && clr != 0;
                wave_lds_sync();
                if (reach2) {
#pragma unroll
                    for (int j = 0; j < 5; j++) {
                        atomicAnd(&ism.seen[q[j] >> 5], ~mk[j]);
                        atomicAnd(&ism.twice[q[j] >> 5], ~mk[j]);
                    }
                }
                if (act) {
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        ism.own[p[j]] = 0xFFFFFFFFu;
                }
                wave_lds_sync();
                if (replay) {
                    for (uint32_t cc = s; cc < cnt; cc++) {
                        uint32_t h[5];
#pragma unroll
                        for (int j = 0; j < 5; j++)
                            h[j] = (uint32_t)__builtin_amdgcn_readlane((int)myh[j], (int)cc);
                        uint32_t nb = 0;
#pragma unroll
                        for (int j = 0; j < 5; j++) {
                            const uint32_t pos = h[j] & 0x7FFu;
                            nb += (fw[pos >> 5] >> (pos & 31)) & 1u ? 0u : 1u;
                            wave_lds_sync();
                            if (lane == 0)
                                fw[pos >> 5] |= 1u << (pos & 31);
                            wave_lds_sync();
                        }
                        if (nb == 0)
                            continue;
                        if constexpr (INDEXED) {
                            // bloom_filter::insert_sha1 on the index: lane j
                            // sets position j; the insert counts when one of
                            // them was clear
                            bool was = false;
                            if (lane < 5) {
                                const uint32_t hp = (lane == 0 ? h[0] : lane == 1 ? h[1] : lane == 2 ? h[2]
                                                     : lane == 3 ? h[3] : h[4]) & imask;
                                const uint32_t bit = 1u << (hp & 31);
                                was = (atomicOr(&ixw[hp >> 5], bit) & bit) != 0;
                            }
                            if (__popcll(__ballot(was)) == 5)
                                continue;
                            ecount++;
                        }
                        uint32_t setc = 0;
#pragma unroll
                        for (int j = 0; j < 5; j++) {
                            const uint32_t bp = h[j] & 0x1FFFFu;
                            setc += (big[bp >> 5] >> (bp & 31)) & 1u;
                            wave_lds_sync();
                            if (lane == 0)
                                big[bp >> 5] |= 1u << (bp & 31);
                            wave_lds_sync();
                        }
                        if (setc == 5)
                            continue;
                        counted(1);
                    }
                    s = cnt;
                    continue;
                }
                // 3. commit up to the candidate that completes a filter
                const unsigned long long am = __ballot(acc);
                const uint32_t need = min(SDBFX_MAX_ELEM - last, SDBFX_BIG_ELEM - bigfi);
                uint32_t cut = cnt - 1;
                if ((uint32_t)__popcll(am) >= need)
                    cut = select_bit(am, need - 1);
                const bool com = act && (uint32_t)lane <= cut;
                if (com) {
#pragma unroll
                    for (int j = 0; j < 5; j++)
                        atomicOr(&fw[p[j] >> 5], 1u << (p[j] & 31));
                    if (acc) {
#pragma unroll
                        for (int j = 0; j < 5; j++)
                            atomicOr(&big[q[j] >> 5], 1u << (q[j] & 31));
                    }
                }
                if constexpr (INDEXED) {
                    // every committed candidate with new filter bits sets its
                    // index bits; those with a clear one count as elements
                    if (com && reach) {
#pragma unroll
                        for (int j = 0; j < 5; j++)
                            if ((iclr >> j) & 1)
                                atomicOr(&ixw[ip[j] >> 5], 1u << (ip[j] & 31));
                    }
                    ecount += (unsigned long long)__popcll(__ballot(com && reach2));
                }
                wave_lds_sync();
                const unsigned long long upto = cut >= 63 ? ~0ull : ((2ull << cut) - 1);
                const uint32_t k = (uint32_t)__popcll(am & upto);
                if (k)
                    counted(k);
                s = cut + 1;
            }
            t = tn;
            c0 = c0n;
            C = Cn;
        }
    }
    flush(bfc - 1);
    if (bfc > 1 && last < SDBFX_MAX_ELEM / 8) {
        bfc--;
        last = SDBFX_MAX_ELEM;
    }
    if (lane == 0) {
        out_counts[3 * fi] = bfc;
        out_counts[3 * fi + 1] = last;
        out_counts[3 * fi + 2] = overflow;
    }
    __syncthreads();
    }
    if constexpr (INDEXED) {
        if (lane == 0)
            *ielem = ecount;
    }
}

__global__ __launch_bounds__(64) void stream_insert_kernel(
    const StreamFile* __restrict__ files, const TileDesc* __restrict__ tiles,
    const TileState* __restrict__ tstate, const uint32_t* __restrict__ cand, uint8_t* __restrict__ stage,
    uint16_t* __restrict__ stage_ham, uint32_t* __restrict__ out_counts)
{
    stream_insert_body<false>(files, blockIdx.x, blockIdx.x + 1, tiles, tstate, cand, stage, stage_ham, out_counts,
                              nullptr, 0u, nullptr);
}

// All `nfiles` inputs, in order, on one wave (launch one workgroup).
__global__ __launch_bounds__(64) void stream_insert_indexed_kernel(
    const StreamFile* __restrict__ files, uint32_t nfiles, const TileDesc* __restrict__ tiles,
    const TileState* __restrict__ tstate, const uint32_t* __restrict__ cand, uint8_t* __restrict__ stage,
    uint16_t* __restrict__ stage_ham, uint32_t* __restrict__ out_counts, uint32_t* __restrict__ ixw, uint32_t imask,
    unsigned long long* __restrict__ ielem)
{
    stream_insert_body<true>(files, 0, nfiles, tiles, tstate, cand, stage, stage_ham, out_counts, ixw, imask, ielem);
}

// ------------------------------------------------------------------ index
//
// Block mode with an index (gen_block_hash, sdbf_core.cc:157-177): every
// counted feature of a block is inserted into info->index, or, with searched
// indexes (info->setlist), every fourth one (k % 4 == 0 in counted order) is
// looked up in each.  block_digest_kernel leaves the counted features' SHA-1
// words in feat (192 slots per block); feature k of launch block b has the
// order key b * 192 + k, which follows the reference's order (inputs, then
// blocks, then features).
//
// Insertion, in three passes.  The index bits are an OR, so any order sets
// the same ones; the element count is the number of features that find one
// of their positions clear, which for feature f means: some position of f
// was clear before the call and no earlier feature of the call has it.  With
// first[p] the smallest key of the features that hold a position p clear
// before the call, f counts exactly when first[p] == key(f) for one of its
// positions.  first[] (one word per index bit, ~0 when unused) is filled by
// index_first_kernel, read by index_commit_kernel (which sets the bits and
// counts), and reset by index_reset_kernel.

// one thread per (block, feature slot) of the launch; false past the block's count
__device__ __forceinline__ bool feat_slot(const BlockDesc* __restrict__ blocks, uint32_t nblk,
                                          const uint16_t* __restrict__ ecf, uint64_t t)
{
    const uint64_t b = t / SDBFX_MAX_ELEM_DD;
    return b < nblk && (t % SDBFX_MAX_ELEM_DD) < ecf[blocks[b].out];
}

__global__ void index_first_kernel(const BlockDesc* __restrict__ blocks, uint32_t nblk,
                                   const uint16_t* __restrict__ ecf, const uint32_t* __restrict__ feat,
                                   const uint32_t* __restrict__ ixw, uint32_t imask, uint32_t* __restrict__ first)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (!feat_slot(blocks, nblk, ecf, t))
        return;
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const uint32_t pp = feat[t * 5 + j] & imask;
        if (!((ixw[pp >> 5] >> (pp & 31)) & 1u))
            atomicMin(&first[pp], (uint32_t)t);
    }
}

__global__ void index_commit_kernel(const BlockDesc* __restrict__ blocks, uint32_t nblk,
                                    const uint16_t* __restrict__ ecf, const uint32_t* __restrict__ feat,
                                    uint32_t* __restrict__ ixw, uint32_t imask, const uint32_t* __restrict__ first,
                                    unsigned long long* __restrict__ ielem)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const bool ok = feat_slot(blocks, nblk, ecf, t);
    uint32_t pp[5];
    bool mine = false;
    if (ok) {
#pragma unroll
        for (int j = 0; j < 5; j++) {
            pp[j] = feat[t * 5 + j] & imask;
            mine |= first[pp[j]] == (uint32_t)t;
        }
        if (mine) {
#pragma unroll
            for (int j = 0; j < 5; j++)
                atomicOr(&ixw[pp[j] >> 5], 1u << (pp[j] & 31));
        }
    }
    const unsigned long long m = __ballot(mine);
    if (m && (threadIdx.x & 63) == (uint32_t)__builtin_ctzll(m))
        atomicAdd(ielem, (unsigned long long)__popcll(m));
}

__global__ void index_reset_kernel(const BlockDesc* __restrict__ blocks, uint32_t nblk,
                                   const uint16_t* __restrict__ ecf, const uint32_t* __restrict__ feat,
                                   uint32_t imask, uint32_t* __restrict__ first)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (!feat_slot(blocks, nblk, ecf, t))
        return;
#pragma unroll
    for (int j = 0; j < 5; j++)
        first[feat[t * 5 + j] & imask] = 0xFFFFFFFFu;
}

// fills n words with v (grid-stride)
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// A searched index (bloom_filter::query_sha1, bloom_filter.cc:334-359): its
// words and bit mask.
struct IndexRef {
    const uint32_t* words;
    uint32_t mask;
    uint32_t pad;
};

// check_indexes (sdbf_core.cc:167-177, 482-493): counted features k % 4 == 0
// of each block looked up in every searched index; matches[b * nix + m]
// counts block b's hits in index m.  One thread per (block, looked-up
// feature).
__global__ void index_search_kernel(const BlockDesc* __restrict__ blocks, uint32_t nblk,
                                    const uint16_t* __restrict__ ecf, const uint32_t* __restrict__ feat,
                                    const IndexRef* __restrict__ ixs, uint32_t nix, uint32_t* __restrict__ matches)
{
    constexpr uint32_t PER = SDBFX_MAX_ELEM_DD / 4;
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t b = t / PER;
    const uint32_t k = 4 * (uint32_t)(t % PER);
    if (b >= nblk || k >= ecf[blocks[b].out])
        return;
    uint32_t h[5];
#pragma unroll
    for (int j = 0; j < 5; j++)
        h[j] = feat[(b * SDBFX_MAX_ELEM_DD + k) * 5 + j];
    for (uint32_t m = 0; m < nix; m++) {
        const IndexRef ix = ixs[m];
        bool all = true;
#pragma unroll
        for (int j = 0; j < 5; j++) {
            const uint32_t pp = h[j] & ix.mask;
            all = all && ((ix.words[pp >> 5] >> (pp & 31)) & 1u);
        }
        if (all)
            atomicAdd(&matches[b * nix + m], 1u);
    }
}

// Copies staged stream filters into their final slots and writes their
// element counts (160 for every filter but the last, which holds last_count;
// sdbf/sdbf_class.cc:407-415).  One wave per input.
__global__ __launch_bounds__(64) void gather_filters_kernel(
    const uint8_t* __restrict__ stage, const uint16_t* __restrict__ stage_ham,
    const uint64_t* __restrict__ src_first, const uint64_t* __restrict__ dst_first,
    const uint32_t* __restrict__ count, const uint32_t* __restrict__ last_count, uint8_t* __restrict__ out,
    uint16_t* __restrict__ ham, uint16_t* __restrict__ ecf)
{
    const uint32_t f = blockIdx.x;
    const uint64_t s = src_first[f], d = dst_first[f];
    const uint32_t nf = count[f], lc = last_count[f];
    for (uint32_t i = 0; i < nf; i++) {
        ((uint32_t*)(out + (d + i) * SDBFX_BF))[threadIdx.x] =
            ((const uint32_t*)(stage + (s + i) * SDBFX_BF))[threadIdx.x];
        if (threadIdx.x == 0) {
            ham[d + i] = stage_ham[s + i];
            ecf[d + i] = (uint16_t)(i + 1 < nf ? SDBFX_MAX_ELEM : lc);
        }
    }
}

// ------------------------------------------------------------------ compare

// Per reference-filter instance and target chunk: the maximum over the chunk's
// target filters (element count >= 16) of sdbf_max_score's per-pair score,
// stored as an exact fraction (num << 16 | den); 0xFFFFFFFF when no target
// filter qualified (the reference's -1).
// A filter set on the device: filters (64 words each), per-filter element
// counts and Hamming weights, and the column order of its target digests
// (col_order_kernel; read by cmp_mfma_kernel only).
struct SetView {
    const uint32_t* filt;
    const uint16_t* ecf;
    const uint16_t* hamf;
    const uint32_t* cord;   // position -> filter index within its digest
    const uint32_t* ckey;   // position -> element count | Hamming weight << 16
};

// FAST: filters folded to 64 bytes (sdbf::fast), scored as sdbf_max_score's
// fast branch (sdbf/sdbf_core.cc:422-437): cutoff CUTOFFS64[1024 / (s1 + s2)]
// and bf_bitcount_cut_64 (sdbf/bf_utils.cc:79-110), whose only early exit is
// 2 x (count of the first 32 bytes) + 48 < cutoff.
// popcount(x) + acc as one v_bcnt_u32_b32: keeps the per-word counts a chain
// of accumulating bcnts (the compiler otherwise sums them as a tree of
// v_add3, 31 more instructions per pair)
#if defined(HIPEMU)
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc) { return (uint32_t)__popc(x) + acc; }
#else
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc)
{
    uint32_t r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}
#endif

// MODE 0: plain 256-byte filters.  MODE 1 (FAST): the reference filter is
// folded (sdbf::fast): 64 bytes, CUTOFFS64, bf_bitcount_cut_64, with the
// target's filter and Hamming weight as stored, folded or not.  MODE 2
// (MIXED): a plain reference filter against a folded target: all 256 bytes
// and the 256-byte rules, with the target's folded Hamming weight; as in
// sdbf_max_score the denominator (max_est - cut_off) is a 32-bit unsigned
// difference, which can wrap, and 0 makes the fraction infinite.  MIXED
// keeps each maximum as a 64-bit (num << 32 | den) with den 0 for infinity.
template <int MODE>
__device__ __forceinline__ void cmp_body(
    SetView A, SetView B, const CmpGroup* __restrict__ groups, uint32_t ngroups,
    const CmpPair* __restrict__ pairs, uint64_t nwork, void* __restrict__ partial_out, uint16_t* cut)
{
    constexpr bool FAST = MODE == 1;
    constexpr bool WIDE = MODE == 2;
    constexpr int W = FAST ? 16 : 64;   // words of a filter compared
    const int lane = threadIdx.x;
    for (int s = lane; s < 1025; s += 64) {
        const uint32_t idx = FAST ? min(1024u / (uint32_t)max(s, 1), (uint32_t)SDBFX_CUTOFFS_LEN - 1)
                                  : (s >= 28 ? 4096u / (uint32_t)s : 0u);
        cut[s] = FAST ? c_cut64[idx] : c_cut[idx];
    }
    __syncthreads();

    for (uint64_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        // group of this work item
        uint32_t lo = 0, hi = ngroups;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (groups[mid].wi_begin <= w)
                lo = mid;
            else
                hi = mid;
        }
        const CmpGroup g = groups[lo];
        const uint64_t local = w - g.wi_begin;
        const uint32_t b = (uint32_t)(local / g.nch);
        const uint32_t c = (uint32_t)(local % g.nch);
        const uint32_t li = b * 64 + lane;
        const bool valid = li < g.inst_count;
        const uint64_t inst = g.inst_begin + li;

        uint32_t a[W];
        uint32_t s1 = 0, e1 = 0;
        uint32_t p = g.pair_begin;
        if (valid) {
            uint32_t plo = g.pair_begin, phi = g.pair_end;
            while (phi - plo > 1) {
                const uint32_t mid = (plo + phi) >> 1;
                if (pairs[mid].inst_begin <= inst)
                    plo = mid;
                else
                    phi = mid;
            }
            p = plo;
            const CmpPair pr = pairs[p];
            const SetView R = pr.ref_set ? B : A;
            const uint64_t rf = pr.ref_first + (inst - pr.inst_begin);
            s1 = R.ecf[rf];
            e1 = R.hamf[rf];
            const uint4* src = (const uint4*)(R.filt + rf * SDBFX_BFW);
#pragma unroll
            for (int q = 0; q < W / 4; q++) {
                const uint4 v = src[q];
                a[4 * q] = v.x;
                a[4 * q + 1] = v.y;
                a[4 * q + 2] = v.z;
                a[4 * q + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < W; q++)
                a[q] = 0;
        }

        uint32_t best_n = 0, best_d = 0;
        bool have = false;   // WIDE: a qualifying target filter was seen
        const uint64_t j0 = g.tgt_first + (uint64_t)c * g.chunk;
        const uint64_t j1 = min(j0 + g.chunk, g.tgt_first + g.tgt_count);
        const uint32_t* __restrict__ tf = g.tgt_set ? B.filt : A.filt;
        const uint16_t* __restrict__ tec = g.tgt_set ? B.ecf : A.ecf;
        const uint16_t* __restrict__ tham = g.tgt_set ? B.hamf : A.hamf;
        for (uint64_t j = j0; j < j1; j++) {
            const uint32_t s2 = tec[j];
            if (s2 < SDBFX_MIN_ELEM)
                continue;
            const uint32_t e2 = tham[j];
            const uint32_t* bj = tf + j * SDBFX_BFW;
            uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
            for (int q = 0; q < 8; q++)
                c0 = bcnt_acc(a[q] & bj[q], c0);
#pragma unroll
            for (int q = 8; q < 16; q++)
                c1 = bcnt_acc(a[q] & bj[q], c1);
            if (!FAST) {
#pragma unroll
                for (int q = 16; q < W / 2; q++)
                    c2 = bcnt_acc(a[q] & bj[q], c2);
#pragma unroll
                for (int q = W / 2; q < W; q++)
                    c3 = bcnt_acc(a[q] & bj[q], c3);
            }
            const uint32_t sum = s1 + s2;
            uint32_t cutoff, match;
            if (FAST) {
                cutoff = sum <= 1024 ? cut[sum] : (uint32_t)c_cut64[0];
                match = 2 * c0 + 48 >= cutoff ? c0 + c1 : 0u;
            } else {
                const uint32_t p32 = c0, p64 = c0 + c1, p128 = p64 + c2, p256 = p128 + c3;
                cutoff = sum <= 1024 ? cut[sum] : 1250u;
                const bool pass = (8 * p32 + 48 >= cutoff) && (4 * p64 + 48 >= cutoff) &&
                                  (2 * p128 + 48 >= cutoff);
                match = pass ? p256 : 0u;
            }
            uint32_t num = 0, den = 1;
            if (match > cutoff) {
                num = match - cutoff;
                den = min(e1, e2) - cutoff;   // WIDE: may wrap or be 0
            }
            if (WIDE) {
                // num / den > best_n / best_d, den 0 being +infinity
                const bool gt = !have || (best_d != 0 && (den == 0 || (uint64_t)num * best_d > (uint64_t)best_n * den));
                if (gt) {
                    best_n = num;
                    best_d = den;
                }
                have = true;
            } else if (best_d == 0 || num * best_d > best_n * den) {
                best_n = num;
                best_d = den;
            }
        }
        if (valid) {
            const uint64_t slot = g.partial_begin + (uint64_t)li * g.nch + c;
            if (WIDE)
                ((uint64_t*)partial_out)[slot] = have ? ((uint64_t)best_n << 32 | best_d) : ~0ull;
            else
                ((uint32_t*)partial_out)[slot] = best_d ? ((best_n << 16) | best_d) : 0xFFFFFFFFu;
        }
    }
}

__global__ __launch_bounds__(64) void cmp_kernel(
    SetView A, SetView B, const CmpGroup* __restrict__ groups, uint32_t ngroups,
    const CmpPair* __restrict__ pairs, uint64_t nwork, void* __restrict__ partial)
{
    __shared__ uint16_t cut[1025];
    cmp_body<0>(A, B, groups, ngroups, pairs, nwork, partial, cut);
}

__global__ __launch_bounds__(64) void cmp_fast_kernel(
    SetView A, SetView B, const CmpGroup* __restrict__ groups, uint32_t ngroups,
    const CmpPair* __restrict__ pairs, uint64_t nwork, void* __restrict__ partial)
{
    __shared__ uint16_t cut[1025];
    cmp_body<1>(A, B, groups, ngroups, pairs, nwork, partial, cut);
}

__global__ __launch_bounds__(64) void cmp_mixed_kernel(
    SetView A, SetView B, const CmpGroup* __restrict__ groups, uint32_t ngroups,
    const CmpPair* __restrict__ pairs, uint64_t nwork, void* __restrict__ partial)
{
    __shared__ uint16_t cut[1025];
    cmp_body<2>(A, B, groups, ngroups, pairs, nwork, partial, cut);
}

// sdbf::fast (sdbf/sdbf_class.cc:426-439) with bloom_filter::fold(2)
// (sdbf/bloom_filter.cc:263-289): each filter's four 64-byte quarters OR-ed
// into its first 64 bytes (the other 192 bytes stay, so folding again changes
// nothing), and its Hamming weight set to the popcount of those 64 bytes.
// One thread per filter.
__global__ void fold_kernel(uint32_t* __restrict__ filt, uint16_t* __restrict__ hamf, uint64_t f0, uint64_t nf)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= nf)
        return;
    uint4* w = (uint4*)(filt + (f0 + t) * SDBFX_BFW);
    uint4 o[4];
    uint32_t h = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 x = w[q], y = w[q + 4], z = w[q + 8], v = w[q + 12];
        o[q] = make_uint4(x.x | y.x | z.x | v.x, x.y | y.y | z.y | v.y, x.z | y.z | z.z | v.z, x.w | y.w | z.w | v.w);
        h += __popc(o[q].x) + __popc(o[q].y) + __popc(o[q].z) + __popc(o[q].w);
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
        w[q] = o[q];
    hamf[f0 + t] = (uint16_t)h;
}

// ------------------------------------------------------- compare on MFMA
//
// The same per-(instance, chunk) maxima as cmp_kernel, with the AND-popcount of
// 128 x 128 filter pairs per wave computed on the matrix cores: every filter
// bit becomes one FP4 (E2M1) element, 0 or 0.5, and the block-scaled MFMA
// v_mfma_scale_f32_32x32x64_f8f6f4 (both scales 2^1, f32 accumulation; a
// 4 x 4 grid of 32 x 32 blocks) multiplies them, so the 2048-element dot
// product of two expanded filters is their exact AND-popcount.  FP4 is the
// densest operand type of the matrix cores: 64 K elements per lane pair per
// instruction, twice the int8 form at the same issue rate.
//
// Operands: a lane gives 32 elements (4 registers, MfOps below), all 32
// bits of one filter word; lanes 0-31 give a word of the chunk's first half,
// lanes 32-63 the same word of its second half.  Over 4 K-steps the two lane
// halves cover words 8q..8q+3 and 8q+4..8q+7; the reference rows (A) and
// target columns (B) use the same nibble-to-bit assignment, so the product
// sums each bit position exactly once.
//
// Only the full 2048-bit count comes out of the matrix cores.  A pair can
// score above 0 only if that count exceeds its cutoff cut(s1 + s2), so every
// accumulator starts at -(cutoff + 1) of its pair (read from an LDS table
// indexed by the sum of the two filters' size codes, size_code below) and
// ends >= 0 exactly for those pairs; all values are whole numbers (plus
// one half in the FP4 form, so that no sum is 0) below 2^24, exact in f32,
// and the sign bit of each sum is the test.  These pairs are
// listed (MF_CAND_CAP per round) and recomputed exactly on the VALU with the
// early-exit checkpoints of bf_bitcount_cut_256; their fractions are merged
// with a 64-bit LDS max on floor(num * 2^22 / den) (strictly monotone in
// num/den for denominators below 2^11) followed by num << 16 | den.
// Result D of one 32x32 MFMA: lane l, register i holds row
// 8 * (i / 4) + 4 * (l / 32) + i % 4, column l % 32.

typedef int sdbfx_v4i __attribute__((vector_size(16)));
typedef int sdbfx_v8i __attribute__((vector_size(32)));
typedef int sdbfx_v16i __attribute__((vector_size(64)));
typedef float sdbfx_v16f __attribute__((vector_size(64)));

#define MF_B 4                                  // 32-wide blocks per tile side
#define MF_CAND_CAP 704                         // candidates listed per tile

__device__ __forceinline__ uint32_t word_of(const uint4& v, int k)
{
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}

// The operands of v_mfma_scale_f32_32x32x64_f8f6f4 in FP4 (E2M1): each lane
// gives 32 elements in 4 registers.  Register j of word w is (w >> j) &
// 0x11111111: 8 nibbles, bit j + 4k of w in nibble k, each 0 or 0b0001 =
// 0.5, so the four registers of a word hold all its 32 bits once.  With both
// block scales 2^1, a product is 1.0 exactly when both bits are set, and the
// f32 accumulator adds whole numbers (at most 2048) to the initial value, all
// exact.
#define MF_SCALE 128   // E8M0 block scale 2^1

// The two operand forms of cmp_mfma_body.  FP4: the form above, 4 K-steps
// of 64 per 8 filter words.  Int8 (v_mfma_i32_32x32x32_i8, int32
// accumulation): byte planes, plane j of word w is (w >> j) & 0x01010101,
// the bits j, 8+j, 16+j, 24+j of w as bytes; a lane gives 16 bytes (planes
// sh..sh+3 of one word), so a word takes 2 K-steps of 32 and 8 words take 8.
template <bool FP4>
struct MfOps;
template <>
struct MfOps<true> {
    typedef sdbfx_v8i Op;
    typedef sdbfx_v16f Acc;
    typedef float El;
    static constexpr int kSteps = 4;
    static constexpr uint32_t kMask = 0x11111111u;
    __device__ static __forceinline__ Acc mfma(const Op& a, const Op& b, const Acc& c)
    {
        return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 4, 4, 0, MF_SCALE, 0, MF_SCALE);
    }
    __device__ static __forceinline__ uint32_t bits(El v) { return __builtin_bit_cast(uint32_t, v); }
    // an accumulator's initial value x as x + 0.5: the sums are then never
    // 0, and x + count >= 0 exactly when the sign bit is clear, whatever the
    // sign of a zero result would be
    __device__ static __forceinline__ El initial(int x) { return (float)x + 0.5f; }
    // the word and first plane of step t of a chunk's 4 words
    __device__ static __forceinline__ int word(int t) { return t; }
    __device__ static __forceinline__ int shift(int) { return 0; }
};
template <>
struct MfOps<false> {
    typedef sdbfx_v4i Op;
    typedef sdbfx_v16i Acc;
    typedef int El;
    static constexpr int kSteps = 8;
    static constexpr uint32_t kMask = 0x01010101u;
    __device__ static __forceinline__ Acc mfma(const Op& a, const Op& b, const Acc& c)
    {
        return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ uint32_t bits(El v) { return (uint32_t)v; }
    __device__ static __forceinline__ El initial(int x) { return x; }
    __device__ static __forceinline__ int word(int t) { return t >> 1; }
    __device__ static __forceinline__ int shift(int t) { return 4 * (t & 1); }
};

// the operand registers of step t (planes shift(t) .. shift(t) + 3 of w; the
// FP4 form's registers 4-7 are not read and stay 0)
template <bool FP4>
__device__ __forceinline__ typename MfOps<FP4>::Op mf_operand(uint32_t w, int t)
{
    typedef MfOps<FP4> M;
    typename M::Op r = {};
    const int sh = M::shift(t);
#pragma unroll
    for (int j = 0; j < 4; j++)
        r[j] = (int)((w >> (sh + j)) & M::kMask);
    return r;
}

// in_agprs() places a 32x32 block of accumulators in AGPRs at that point,
// so that its initial values are not held in VGPRs up to the loop (plain form
// under the CPU emulation).
#if defined(HIPEMU)
__device__ __forceinline__ void in_agprs(sdbfx_v16f&) {}
__device__ __forceinline__ void in_agprs(sdbfx_v16i&) {}
#else
__device__ __forceinline__ void in_agprs(sdbfx_v16f& v) { asm volatile("" : "+a"(v)); }
__device__ __forceinline__ void in_agprs(sdbfx_v16i& v) { asm volatile("" : "+a"(v)); }
#endif

// Index part of a filter with s elements in the cutoff table cutx: cutx[a + b]
// is the reference's cutoff for two qualifying filters (>= 16 elements), 1250
// for sums above 1024, and 0xFFFF (never exceeded) when either filter has
// fewer than 16 elements.
__device__ __forceinline__ uint32_t size_code(uint32_t s)
{
    return s >= SDBFX_MIN_ELEM ? min(s, 1025u) : 2048u;
}

// AND-popcounts of the first 32, 64, 128 and all 256 bytes of a reference
// filter and a target filter, as bf_bitcount_cut_256 (sdbf/bf_utils.cc:115-206)
// takes them.  The reference filter's 16-byte chunk q is ra[q ^ sw] (the
// swizzled LDS tile).
__device__ __forceinline__ void prefix_counts(const uint4* ra, uint32_t sw, const uint4* __restrict__ fb,
                                              uint32_t& p32, uint32_t& p64, uint32_t& p128, uint32_t& p256)
{
    // chunks 0-1, 2-3, 4-7, 8-15: the target filter's 16 loads issued
    // together (one round trip), the reference side read from LDS four
    // chunks at a time
    uint32_t c[4] = {0, 0, 0, 0};
    uint4 yy[16];
#pragma unroll
    for (int k = 0; k < 16; k++)
        yy[k] = fb[k];
#pragma unroll
    for (int g = 0; g < 4; g++) {
        uint4 x[4], y[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            x[k] = ra[(4 * g + k) ^ sw];
            y[k] = yy[4 * g + k];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int q = 4 * g + k;
            c[q < 2 ? 0 : q < 4 ? 1 : q < 8 ? 2 : 3] +=
                __popc(x[k].x & y[k].x) + __popc(x[k].y & y[k].y) + __popc(x[k].z & y[k].z) + __popc(x[k].w & y[k].w);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    p32 = c[0];
    p64 = p32 + c[1];
    p128 = p64 + c[2];
    p256 = p128 + c[3];
}

// The early-exit checks of bf_bitcount_cut_256 and the fraction of
// sdbf_max_score (sdbf_core.cc:418-440): true with num/den when the pair
// passes and its match exceeds the cutoff, merged into *best as a 64-bit
// key floor(num * 2^22 / den) << 32 | num << 16 | den.
__device__ __forceinline__ void merge_pair(uint32_t p32, uint32_t p64, uint32_t p128, uint32_t p256,
                                           uint32_t cutoff, uint32_t e1, uint32_t e2, unsigned long long* best)
{
    const bool pass = (8 * p32 + 48 >= cutoff) && (4 * p64 + 48 >= cutoff) && (2 * p128 + 48 >= cutoff);
    const uint32_t match = pass ? p256 : 0u;
    if (match <= cutoff)
        return;
    const uint32_t num = match - cutoff, den = min(e1, e2) - cutoff;
    const unsigned long long key = ((unsigned long long)num << 22) / den;
    atomicMax(best, key << 32 | (unsigned long long)(num << 16 | den));
}

// Per-wave LDS of cmp_mfma_kernel.
struct MfmaWaveSmem {
    // the 128 reference filters, 16-byte chunk c of row r at r * 16 + (c ^ (r & 15))
    // so that the 32 rows read by one instruction fall in different banks
    uint4 atile[SDBFX_MROWS * 16];
    uint32_t rinfo[SDBFX_MROWS];                    // size_code(s1) | e1 << 16
    uint64_t rfil[SDBFX_MROWS];                     // reference filter of each row: index | set << 63
    unsigned long long best[SDBFX_MROWS];
    uint32_t clist[MF_CAND_CAP];    // row | (target filter index in its digest) << 7
    uint32_t bfix[SDBFX_MROWS];     // per row: floor(best fraction * 2^16), or its initial accumulator
};

// Column order of a target digest for cmp_mfma_kernel.  A 128-column tile
// whose columns share one size code takes per-row initial values (a quarter
// of the tile's set-up otherwise), so the filters of every code held by at
// least 128 filters of the digest are sorted by code, one tile after another.
// The filters of rarer codes (among them the low element counts, whose low
// cutoffs let many unrelated pairs through) stay mixed, as in the digest: met
// as a run of 128, they would be compared with rows whose best fraction is
// still low, and most of their pairs would become candidates.  The rare ones
// come first, then the sorted ones; the 128-filter tiles are then dealt over
// the digest's positions with a golden-ratio stride (position tile P holds
// tile P * k mod m of that order, m full tiles, k coprime with m; a partial
// last tile stays last), so that every range of positions that a work item
// takes mixes all of them.  A counting sort, one workgroup per digest.  The
// maximum over a digest's filters does not depend on their order, so no
// score changes.  Position p of the digest (from its first filter f0) holds
// cord[f0 + p], the index of a filter within the digest, and ckey[f0 + p], its
// element count | Hamming weight << 16.
#define CO_BINS 1024
__device__ __forceinline__ uint32_t co_bin(uint32_t s)
{
    return s >= SDBFX_MIN_ELEM ? min(s, 1025u) - SDBFX_MIN_ELEM : CO_BINS - 1;
}

// k^-1 mod m for k near 0.618 m with gcd(k, m) = 1 (1 when m <= 2).
__device__ uint32_t co_stride_inverse(uint32_t m)
{
    if (m <= 2)
        return 1;
    uint32_t k = max(1u, (uint32_t)(((uint64_t)m * 2654435769ull) >> 32));
    for (;; k++) {
        // extended Euclid on (m, k): gcd and the inverse of k
        int64_t t0 = 0, t1 = 1, r0 = m, r1 = k;
        while (r1) {
            const int64_t q = r0 / r1, t2 = t0 - q * t1, r2 = r0 - q * r1;
            t0 = t1;
            t1 = t2;
            r0 = r1;
            r1 = r2;
        }
        if (r0 == 1)
            return (uint32_t)(t0 < 0 ? t0 + m : t0);
    }
}

// Adds 1 to ctr[key] for every lane with `on` and returns the lane's place
// (the counter's value before its increment).  Most filters of a digest share
// one element count, so the lanes with lane 0's key (or the first active
// lane's) take one atomic together; the others take one each.
__device__ __forceinline__ uint32_t lds_inc_lead(uint32_t* ctr, uint32_t key, bool on)
{
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long act = __ballot(on);
    if (!act)
        return 0;
    const int leader = __builtin_ctzll(act);
    const uint32_t lkey = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
    const unsigned long long same = __ballot(on && key == lkey);
    uint32_t base = 0;
    if ((int)lane == leader)
        base = atomicAdd(&ctr[lkey], (uint32_t)__popcll(same));
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    if ((same >> lane) & 1)
        return base + lanes_below(same);
    return on ? atomicAdd(&ctr[key], 1u) : 0u;
}

__device__ __forceinline__ void col_order_body(const uint16_t* __restrict__ ecf, const uint16_t* __restrict__ hamf,
                                               const uint64_t* __restrict__ dfirst,
                                               const uint32_t* __restrict__ dcount, uint32_t* __restrict__ cord,
                                               uint32_t* __restrict__ ckey, uint32_t d)
{
    constexpr uint32_t kRare = 0x80000000u;   // bin start of a rare code
    constexpr uint32_t kPer = 8;              // filters per thread per step, loaded first
    // per bin: its count, then its next rank; [CO_BINS]: the next rank of the rare codes
    __shared__ uint32_t hist[CO_BINS + 1];
    __shared__ uint32_t wsum[16], wsumr[16];
    __shared__ uint32_t kinv;
    const uint64_t f0 = dfirst[d];
    const uint32_t n = dcount[d];
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const uint32_t m = n / 128;   // full tiles
    hist[t] = 0;
    if (t == 0) {
        kinv = m ? co_stride_inverse(m) : 1u;
        hist[CO_BINS] = 0;
    }
    __syncthreads();
    for (uint32_t i0 = 0; i0 < n; i0 += kPer * 1024) {
        uint32_t e[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t i = i0 + k * 1024 + t;
            e[k] = i < n ? ecf[f0 + i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++)
            (void)lds_inc_lead(hist, co_bin(e[k]), i0 + k * 1024 + t < n);
    }
    __syncthreads();
    // bin starts: rare codes share ranks 0 .. R-1 (hist[CO_BINS]), the others
    // follow in code order.  Thread t owns bin t.
    const uint32_t v = hist[t];
    uint32_t wtot, wrare;
    uint32_t run = wave_excl_sum(v >= 128 ? v : 0u, (int)lane, &wtot);
    (void)wave_excl_sum(v >= 128 ? 0u : v, (int)lane, &wrare);
    if (lane == 0) {
        wsum[wv] = wtot;
        wsumr[wv] = wrare;
    }
    __syncthreads();
    uint32_t nrare = 0;
    for (uint32_t q = 0; q < 16; q++)
        nrare += wsumr[q];
    for (uint32_t q = 0; q < wv; q++)
        run += wsum[q];
    hist[t] = v >= 128 ? nrare + run : kRare;
    __syncthreads();
    const uint32_t ki = kinv;
    for (uint32_t i0 = 0; i0 < n; i0 += kPer * 1024) {
        uint32_t e[kPer], hm[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t i = i0 + k * 1024 + t;
            e[k] = i < n ? ecf[f0 + i] : 0u;
            hm[k] = i < n ? hamf[f0 + i] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t i = i0 + k * 1024 + t;
            const uint32_t bin = co_bin(e[k]);
            // bins of rare codes keep kRare; the counters stay below 2^31
            const uint32_t r = lds_inc_lead(hist, (hist[bin] & kRare) ? (uint32_t)CO_BINS : bin, i < n);
            if (i < n) {
                const uint32_t p = r < m * 128 ? (uint32_t)((uint64_t)(r / 128) * ki % m) * 128 + r % 128 : r;
                cord[f0 + p] = i;
                ckey[f0 + p] = e[k] | hm[k] << 16;
            }
        }
    }
}

__global__ __launch_bounds__(1024) void col_order_kernel(const uint16_t* __restrict__ ecf,
                                                         const uint16_t* __restrict__ hamf,
                                                         const uint64_t* __restrict__ dfirst,
                                                         const uint32_t* __restrict__ dcount,
                                                         uint32_t* __restrict__ cord, uint32_t* __restrict__ ckey)
{
    col_order_body(ecf, hamf, dfirst, dcount, cord, ckey, blockIdx.x);
}

// The set-up of cmp_mfma_kernel in one launch.  Workgroups [0, nd): the column
// order of nd target digests (col_order_body).  The rest: one instance per
// thread, its row data (where its reference filter is, its size code and
// Hamming weight) and its shared maximum gbest cleared.
__global__ __launch_bounds__(1024) void cmp_prep_kernel(const uint16_t* __restrict__ ecf,
                                                        const uint16_t* __restrict__ hamf,
                                                        const uint64_t* __restrict__ dfirst,
                                                        const uint32_t* __restrict__ dcount, uint32_t nd,
                                                        uint32_t* __restrict__ cord, uint32_t* __restrict__ ckey,
                                                        SetView A, SetView B, const CmpPair* __restrict__ pairs,
                                                        uint32_t npairs, uint64_t ninst, uint64_t* __restrict__ rmap,
                                                        uint32_t* __restrict__ rcode,
                                                        unsigned long long* __restrict__ gbest)
{
    if (blockIdx.x < nd) {
        col_order_body(ecf, hamf, dfirst, dcount, cord, ckey, blockIdx.x);
        return;
    }
    const uint64_t t = (uint64_t)(blockIdx.x - nd) * blockDim.x + threadIdx.x;
    if (t >= ninst)
        return;
    uint32_t lo = 0, hi = npairs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pairs[mid].inst_begin <= t)
            lo = mid;
        else
            hi = mid;
    }
    const CmpPair P = pairs[lo];
    const SetView R = P.ref_set ? B : A;
    const uint64_t rf = P.ref_first + (t - P.inst_begin);
    rmap[t] = rf | (uint64_t)P.ref_set << 63;
    rcode[t] = size_code(R.ecf[rf]) | (uint32_t)R.hamf[rf] << 16;
    gbest[t] = 0ull;
}

// Four waves per workgroup, one per SIMD, each taking work items on its own;
// they share only the cutoff table.
//
// Pruning.  Only the maximum fraction of each reference filter (row) is
// needed.  A pair (i, j) can raise row i's maximum b only if its
// match - cutoff exceeds b * (min(e1, e2) - cutoff), so its accumulator
// starts m = floor(B * (min(e1, e2) - cutoff) / 2^16) lower, with
// B = floor(b * 2^16): only pairs with match - cutoff > m become candidates,
// and every pair that could raise the maximum is one.  b is the row's best so
// far: this work item's, seeded from gbest, the per-instance maximum that all
// work items of the call publish when they finish (any value there is a real
// pair's fraction of the same instance and target digest, so taking it into
// this chunk's partial maximum changes no result).
//
// Latency.  The next tile's column data (element counts, Hamming weights) is
// loaded before a tile's K loop and used after it.
template <bool FP4>
__device__ __forceinline__ void cmp_mfma_body(
    SetView A, SetView B, const CmpGroup* __restrict__ groups, uint32_t ngroups,
    const uint64_t* __restrict__ rmap, const uint32_t* __restrict__ rcode, uint64_t nwork,
    uint32_t* __restrict__ partial, unsigned long long* __restrict__ gbest)
{
    typedef MfOps<FP4> M;
    typedef typename M::El El;
    __shared__ int16_t ncut[4097];                  // -(cutoff + 1) by size_code(s1) + size_code(s2)
    __shared__ MfmaWaveSmem wsm[4];

    for (int s = threadIdx.x; s < 4097; s += 256) {
        const uint32_t idx = s >= 28 ? 4096u / (uint32_t)s : 0u;
        ncut[s] = s > 2050 ? (int16_t)-32768 : s > 1024 ? (int16_t)-1251 : (int16_t)(-1 - (int)c_cut[idx]);
    }
    __syncthreads();

    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int h = lane >> 5, r32 = lane & 31;
    uint4* const atile = wsm[wv].atile;
    uint32_t* const rinfo = wsm[wv].rinfo;
    uint64_t* const rfil = wsm[wv].rfil;
    unsigned long long* const best = wsm[wv].best;
    uint32_t* const clist = wsm[wv].clist;
    uint32_t* const bfix = wsm[wv].bfix;
    const char* const ncutb = (const char*)ncut;

    // each XCD (workgroups are dealt round-robin over 8) takes a contiguous
    // range of work items, so neighbouring items share its L2
    const uint32_t xcd = blockIdx.x & 7, nslot = (gridDim.x >> 3) * 4, slot = (blockIdx.x >> 3) * 4 + wv;
    const uint64_t w_lo = nwork * xcd / 8, w_hi = nwork * (xcd + 1) / 8;

    for (uint64_t w = w_lo + slot; w < w_hi; w += nslot) {
        uint32_t lo = 0, hi = ngroups;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (groups[mid].wi_begin <= w)
                lo = mid;
            else
                hi = mid;
        }
        const CmpGroup g = groups[lo];
        // chunk-major: the items of one row block are nblk apart, so its
        // later chunks start from the maxima its earlier ones published
        const uint64_t local = w - g.wi_begin;
        const uint32_t c = (uint32_t)(local / g.nblk);
        const uint32_t b = (uint32_t)(local % g.nblk);

        // rows: instance b*128 + 32*bm + r32 (lanes r32 and r32+32 share it)
        if (h == 0) {
#pragma unroll
            for (int bm = 0; bm < MF_B; bm++) {
                const uint32_t row = 32 * bm + r32;
                const uint32_t li = b * SDBFX_MROWS + row;
                const bool valid = li < g.inst_count;
                const uint64_t inst = g.inst_begin + (valid ? li : g.inst_count - 1);
                rfil[row] = rmap[inst];
                rinfo[row] = valid ? rcode[inst] : size_code(0);
                best[row] = valid ? gbest[inst] : 0ull;
            }
        }
        wave_lds_sync();
        // 16 loads in flight per lane, then their LDS stores.  The filter
        // address is formed from the kernel's set arguments (not kept in LDS),
        // so that the loads are global ones.
#pragma unroll
        for (uint32_t k0 = 0; k0 < SDBFX_MROWS * 16; k0 += 64 * 16) {
            uint4 v[16];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t k = k0 + lane + 64 * i;
                const uint64_t rm = rfil[k >> 4];
                const uint32_t* __restrict__ rf = ((rm >> 63) ? B.filt : A.filt) + (rm & ~(1ull << 63)) * SDBFX_BFW;
                v[i] = ((const uint4*)rf)[k & 15];
            }
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t k = k0 + lane + 64 * i, r = k >> 4, ch = k & 15;
                atile[r * 16 + (ch ^ (r & 15))] = v[i];
            }
        }
        wave_lds_sync();

        const uint64_t j0 = g.tgt_first + (uint64_t)c * g.chunk;
        const uint32_t nj = (uint32_t)min((uint64_t)g.chunk, g.tgt_count - (uint64_t)c * g.chunk);
        const uint32_t* __restrict__ tf = g.tgt_set ? B.filt : A.filt;
        const uint32_t* __restrict__ tord = g.tgt_set ? B.cord : A.cord;
        const uint32_t* __restrict__ tkey = g.tgt_set ? B.ckey : A.ckey;
        const uint16_t* __restrict__ tec = (g.tgt_set ? B.ecf : A.ecf) + g.tgt_first;
        const uint16_t* __restrict__ tham = (g.tgt_set ? B.hamf : A.hamf) + g.tgt_first;
        // the target digest's filters as 16-byte chunks: filter L of the
        // digest, chunk q at L * 16 + q
        const uint4* __restrict__ tbase = (const uint4*)(tf + g.tgt_first * SDBFX_BFW);
        bool anyq = false;

        // column data of a tile in column order: the filter (index in the
        // digest; columns past the end repeat the last one), and its element
        // count | Hamming weight << 16 (0 past the end)
        auto col_load = [&](uint32_t t0, uint32_t* cl, uint32_t* ky) {
            const uint32_t ncol = min(128u, nj - t0);
#pragma unroll
            for (int bn = 0; bn < MF_B; bn++) {
                const uint32_t col = 32 * bn + r32;
                const uint64_t p = j0 + t0 + min(col, ncol - 1);
                cl[bn] = tord[p];
                ky[bn] = col < ncol ? tkey[p] : 0u;
            }
        };
        uint32_t cl[MF_B], ky[MF_B];
        col_load(0, cl, ky);
        // words 8q + 4h .. 8q + 4h + 3 of a row or column are its chunk 2q + h;
        // bw: the columns' first chunk, loaded ahead by the tile before when
        // bw_ready
        uint4 bw[MF_B];
        bool bw_ready = false;

        for (uint32_t t0 = 0; t0 < nj; t0 += 32 * MF_B) {
            const uint32_t tn0 = t0 + 32 * MF_B < nj ? t0 + 32 * MF_B : t0;   // the next tile (or this one)
            uint32_t ccode[MF_B], sx = 0, cmn = 0xFFFFu, cmx = 0;
#pragma unroll
            for (int bn = 0; bn < MF_B; bn++) {
                const uint32_t s2 = ky[bn] & 0xFFFFu;
                const bool q = s2 >= SDBFX_MIN_ELEM;
                anyq |= q;
                sx |= q;
                ccode[bn] = 2 * size_code(s2);
                cmn = min(cmn, ccode[bn]);
                cmx = max(cmx, ccode[bn]);
            }
            // each lane's columns as uint4 indices of their first chunk
            uint32_t e2c[MF_B], cbt[MF_B];
#pragma unroll
            for (int bn = 0; bn < MF_B; bn++) {
                e2c[bn] = (ky[bn] >> 16) + 1;
                cbt[bn] = cl[bn] * 16 + h;
            }
            // the next tile's column data, consumed after this tile
            col_load(tn0, cl, ky);
            // no qualifying target filter: no pair of the tile can pass
            if (__ballot(sx != 0) == 0) {
                bw_ready = false;
                continue;
            }
            // all columns of one size code: the cutoff depends on the row only
            const uint32_t umin = (uint32_t)__builtin_amdgcn_readlane((int)wave_prefix_min(cmn, lane), 63);
            const uint32_t umax =
                0xFFFFu - (uint32_t)__builtin_amdgcn_readlane((int)wave_prefix_min(0xFFFFu - cmx, lane), 63);
            const bool uniform = umin == umax;
            uint4 aw[MF_B];
#pragma unroll
            for (int k = 0; k < MF_B; k++) {
                aw[k] = atile[(32 * k + r32) * 16 + (h ^ (r32 & 15))];
                if (!bw_ready)
                    bw[k] = tbase[cbt[k]];
            }

            // the smallest qualifying Hamming weight (+1) of the tile
            uint32_t emn = 0xFFFFu;
#pragma unroll
            for (int bn = 0; bn < MF_B; bn++)
                emn = min(emn, ccode[bn] < 4096u ? e2c[bn] : 0xFFFFu);
            const uint32_t e2min = (uint32_t)__builtin_amdgcn_readlane((int)wave_prefix_min(emn, lane), 63);
            // each row's best fraction so far as B = floor(b * 2^16) (two rows
            // per lane).  When all columns share one size code, the row's
            // initial accumulator itself: -(cutoff + 1) - m with the margin
            // from the smallest Hamming weight of the tile, which bounds every
            // pair's margin from below.
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const uint32_t row = lane + 64 * k;
                const uint32_t key = (uint32_t)best[row];
                const uint32_t B = key ? ((key >> 16) << 16) / (key & 0xFFFFu) : 0u;
                if (uniform) {
                    const uint32_t x = rinfo[row];
                    const int n = *(const int16_t*)(ncutb + 2 * (x & 0xFFFFu) + umin);
                    const int d = (int)min((x >> 16) + 1, e2min) + n;
                    bfix[row] = (uint32_t)(n - (int)((B * (uint32_t)max(d, 0)) >> 16));
                } else {
                    bfix[row] = B;
                }
            }
            wave_lds_sync();

            // accumulators start at -(cutoff + 1) - m of their pair, so that
            // after the 2048 bits they are >= 0 exactly when
            // count - cutoff > m.  m is computed with a 24-bit multiply; when
            // min(e1, e2) <= cutoff (d <= 0) the pair cannot pass whatever m
            // is, and the wrapped product still gives m >= 0.
            typename M::Acc acc[MF_B][MF_B];
            uint32_t rb = h;
            opaque(rb);
#pragma unroll
            for (int bm = 0; bm < MF_B; bm++) {
                if (uniform) {
                    El v[16];
#pragma unroll
                    for (int i = 0; i < 16; i++)
                        v[i] = M::initial((int)bfix[4 * rb + 32 * bm + 8 * (i >> 2) + (i & 3)]);
#pragma unroll
                    for (int bn = 0; bn < MF_B; bn++) {
#pragma unroll
                        for (int i = 0; i < 16; i++)
                            acc[bm][bn][i] = v[i];
                        in_agprs(acc[bm][bn]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                } else {
                    uint32_t rc[16], e1c[16], bq[16];
#pragma unroll
                    for (int i = 0; i < 16; i++) {
                        const uint32_t row = 4 * rb + 32 * bm + 8 * (i >> 2) + (i & 3);
                        const uint32_t x = rinfo[row];
                        rc[i] = 2 * (x & 0xFFFFu);
                        e1c[i] = (x >> 16) + 1;
                        bq[i] = bfix[row];
                    }
#pragma unroll
                    for (int bn = 0; bn < MF_B; bn++) {
#pragma unroll
                        for (int i = 0; i < 16; i++) {
                            const int n = *(const int16_t*)(ncutb + rc[i] + ccode[bn]);
                            const uint32_t d = min(e1c[i], e2c[bn]) + (uint32_t)n;
                            acc[bm][bn][i] = M::initial(n - (int)(((bq[i] & 0xFFFFFFu) * (d & 0xFFFFFFu)) >> 16));
                        }
                        in_agprs(acc[bm][bn]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
            }

            // operands of the first step
            typename M::Op av[MF_B], bv[MF_B];
#pragma unroll
            for (int k = 0; k < MF_B; k++) {
                av[k] = mf_operand<FP4>(aw[k].x, 0);
                bv[k] = mf_operand<FP4>(bw[k].x, 0);
            }
#pragma unroll 1
            for (int q = 0; q < 8; q++) {
                // prefetch of the next 8 words, consumed one iteration later
                const int qn = q < 7 ? q + 1 : 7;
                uint32_t ao = (uint32_t)(r32 * 16 + ((2 * qn + h) ^ (r32 & 15)));
                opaque(ao);
                uint4 an[MF_B], bnx[MF_B];
#pragma unroll
                for (int k = 0; k < MF_B; k++) {
                    an[k] = atile[ao + 32 * 16 * k];
                    bnx[k] = tbase[cbt[k] + 2 * qn];
                }
                __builtin_amdgcn_sched_barrier(0);
                // step t: word M::word(t) of the chunk, planes M::shift(t) ..
                // M::shift(t) + 3 (FP4: all 32 bits of word t, 64 K elements
                // with the other lane half's word).  Each MFMA of step t is
                // followed by two operand registers (four VALU instructions)
                // of step t + 1, so that they run in its shadow.
#pragma unroll
                for (int t = 0; t < M::kSteps; t++) {
                    const int shn = M::shift(t + 1);
                    uint32_t wn[2 * MF_B];
#pragma unroll
                    for (int k = 0; k < MF_B; k++) {
                        wn[k] = t < M::kSteps - 1 ? word_of(aw[k], M::word(t + 1)) : an[k].x;
                        wn[MF_B + k] = t < M::kSteps - 1 ? word_of(bw[k], M::word(t + 1)) : bnx[k].x;
                    }
                    typename M::Op nv[2 * MF_B];
#pragma unroll
                    for (int k = 0; k < 2 * MF_B; k++)
                        nv[k] = typename M::Op{};
#pragma unroll
                    for (int gi = 0; gi < MF_B * MF_B; gi++) {
                        const int bm = gi / MF_B, bn = gi % MF_B;
                        acc[bm][bn] = M::mfma(av[bm], bv[bn], acc[bm][bn]);
#pragma unroll
                        for (int e2i = 0; e2i < 2; e2i++) {
                            const int el = 2 * gi + e2i, p = el >> 2, e = el & 3;
                            nv[p][e] = (int)((wn[p] >> (shn + e)) & M::kMask);
                        }
                        __builtin_amdgcn_sched_barrier(0);
                    }
#pragma unroll
                    for (int k = 0; k < MF_B; k++) {
                        av[k] = nv[k];
                        bv[k] = nv[MF_B + k];
                    }
                }
#pragma unroll
                for (int k = 0; k < MF_B; k++) {
                    aw[k] = an[k];
                    bw[k] = bnx[k];
                }
            }
            // the next tile's first column chunks (cl holds its columns), in
            // flight during the candidate pass
#pragma unroll
            for (int k = 0; k < MF_B; k++)
                bw[k] = tbase[cl[k] * 16 + h];
            bw_ready = true;

            // candidates: accumulators >= 0, one bit per accumulator, from
            // the sign bits (shifted in by v_alignbit).  Word w holds blocks
            // 2w (bits 31..16, register i at bit 31 - i) and 2w + 1 (bits
            // 15..0, register i at bit 15 - i).
            // One block of 16 accumulators at a time (the scheduling
            // barriers keep the compiler from reading all 256 at once).
            uint32_t cbits[MF_B * MF_B / 2];
            uint32_t anyc = 0;
#pragma unroll
            for (int wi = 0; wi < MF_B * MF_B / 2; wi++) {
                uint32_t x = 0;
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const int bm = (2 * wi + k) / MF_B, bn = (2 * wi + k) % MF_B;
                    // a new value of the block, defined here in AGPRs, so that
                    // its copies to VGPRs are made here and not all at once
                    in_agprs(acc[bm][bn]);
                    // (each element copied out first: a bit cast of the
                    // subscripted vector itself reads element 0 for every i)
#pragma unroll
                    for (int i = 0; i < 16; i++) {
                        const El v = acc[bm][bn][i];
                        x = __builtin_amdgcn_alignbit(x, M::bits(v), 31);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
                cbits[wi] = ~x;
                anyc |= ~x;
            }
            if (__ballot(anyc != 0) == 0)
                continue;
            // each lane lists its own candidates after those of lower lanes;
            // MF_CAND_CAP of them per round
            uint32_t own = 0;
#pragma unroll
            for (int wi = 0; wi < MF_B * MF_B / 2; wi++)
                own += (uint32_t)__popc(cbits[wi]);
            uint32_t ncand;
            const uint32_t first = wave_excl_sum(own, lane, &ncand);
            for (uint32_t round = 0; round < ncand; round += MF_CAND_CAP) {
                uint32_t pos = first;
                if (pos < round + MF_CAND_CAP && pos + own > round) {
#pragma unroll
                    for (int wi = 0; wi < MF_B * MF_B / 2; wi++) {
                        uint32_t x = cbits[wi];
                        while (x) {
                            const uint32_t bit = 31u - (uint32_t)__clz(x);
                            x ^= 1u << bit;
                            if (pos >= round && pos < round + MF_CAND_CAP) {
                                const uint32_t blk = 2 * wi + (bit < 16 ? 1 : 0), i = (bit < 16 ? 15u : 31u) - bit;
                                // the column's filter: this lane's, block column blk % MF_B
                                const uint32_t cb = bit < 16 ? cbt[(2 * wi + 1) % MF_B] : cbt[(2 * wi) % MF_B];
                                const uint32_t row = 32 * (blk / MF_B) + 8 * (i >> 2) + 4 * h + (i & 3);
                                clist[pos - round] = row | (cb >> 4) << 7;
                            }
                            pos++;
                        }
                    }
                }
                wave_lds_sync();
                // one candidate per lane
                const uint32_t nlist = min(ncand - round, (uint32_t)MF_CAND_CAP);
                for (uint32_t base = 0; base < nlist; base += 64) {
                    if (base + lane < nlist) {
                        const uint32_t e = clist[base + lane];
                        const uint32_t row = e & 127u, L = e >> 7;
                        const uint32_t s2 = tec[L], e2 = tham[L];
                        const uint32_t ri = rinfo[row];
                        uint32_t p32, p64, p128, p256;
                        prefix_counts(atile + row * 16, row & 15, tbase + (uint64_t)L * 16, p32, p64, p128, p256);
                        const uint32_t cutoff = (uint32_t)(-1 - (int)ncut[(ri & 0xFFFFu) + size_code(s2)]);
                        merge_pair(p32, p64, p128, p256, cutoff, ri >> 16, e2, &best[row]);
                    }
                }
                wave_lds_sync();
            }
        }

        const bool hq = __ballot(anyq) != 0;
        if (h == 0) {
#pragma unroll
            for (int bm = 0; bm < MF_B; bm++) {
                const uint32_t row = 32 * bm + r32;
                const uint32_t li = b * SDBFX_MROWS + row;
                if (li < g.inst_count) {
                    const unsigned long long v = best[row];
                    partial[g.partial_begin + (uint64_t)li * g.nch + c] = !hq ? 0xFFFFFFFFu : v ? (uint32_t)v : 1u;
                    if (v)
                        atomicMax(&gbest[g.inst_begin + li], v);
                }
            }
        }
        wave_lds_sync();
    }
}

// The matrix-core comparison with FP4 operands (the default form) and with
// int8 byte planes; both give the same partial maxima.
__global__ __launch_bounds__(256) void cmp_mfma_kernel(
    SetView A, SetView B, const CmpGroup* __restrict__ groups, uint32_t ngroups,
    const uint64_t* __restrict__ rmap, const uint32_t* __restrict__ rcode, uint64_t nwork,
    uint32_t* __restrict__ partial, unsigned long long* __restrict__ gbest)
{
    cmp_mfma_body<true>(A, B, groups, ngroups, rmap, rcode, nwork, partial, gbest);
}

__global__ __launch_bounds__(256) void cmp_mfma_i8_kernel(
    SetView A, SetView B, const CmpGroup* __restrict__ groups, uint32_t ngroups,
    const uint64_t* __restrict__ rmap, const uint32_t* __restrict__ rcode, uint64_t nwork,
    uint32_t* __restrict__ partial, unsigned long long* __restrict__ gbest)
{
    cmp_mfma_body<false>(A, B, groups, ngroups, rmap, rcode, nwork, partial, gbest);
}

// Per reference-filter instance: the largest fraction over its target
// chunks (sdbf_max_score's maximum, sdbf_core.cc:396-455, over all of the
// target digest), or 0xFFFFFFFF when no chunk had a qualifying target.  One
// thread per instance; maxima are order-free, so this part is parallel.
__global__ void chunk_max_kernel(const CmpPair* __restrict__ pairs, uint32_t npairs, uint64_t ninst,
                                 const uint32_t* __restrict__ partial, uint32_t* __restrict__ imax)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= ninst)
        return;
    uint32_t lo = 0, hi = npairs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pairs[mid].inst_begin <= t)
            lo = mid;
        else
            hi = mid;
    }
    const CmpPair P = pairs[lo];
    const uint32_t* v = partial + P.partial_begin + (t - P.inst_begin) * P.nch;
    uint32_t bn = 0, bd = 0;
    for (uint32_t c = 0; c < P.nch; c++) {
        const uint32_t x = v[c];
        if (x == 0xFFFFFFFFu)
            continue;
        const uint32_t n = x >> 16, d = x & 0xFFFFu;
        if (bd == 0 || n * bd > bn * d) {
            bn = n;
            bd = d;
        }
    }
    imax[t] = bd ? (bn << 16 | bd) : 0xFFFFFFFFu;
}

// sdbf_score (sdbf_core.cc:331-389) for each oriented pair: ordered double
// sum of the per-filter maxima with the reference's -1 sentinel handling,
// sparse-filter denominator and floor(x + 0.5) rounding.
__global__ void pair_sum_kernel(const CmpPair* __restrict__ pairs, uint32_t npairs,
                                const uint16_t* __restrict__ ecfA, const uint16_t* __restrict__ ecfB,
                                const uint32_t* __restrict__ imax, int32_t* __restrict__ scores)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs)
        return;
    const CmpPair P = pairs[p];
    const uint16_t* __restrict__ ecf = P.ref_set ? ecfB : ecfA;
    double sum = -1.0;
    uint32_t sparse = 0;
    for (uint32_t i = 0; i < P.ref_count; i++) {
        const uint32_t s1 = ecf[P.ref_first + i];
        double m;
        if (s1 < SDBFX_MIN_ELEM) {
            m = 0.0;
            sparse++;
        } else {
            const uint32_t x = imax[P.inst_begin + i];
            m = x == 0xFFFFFFFFu ? -1.0 : (double)(x >> 16) / (double)(x & 0xFFFFu);
        }
        sum = (sum < 0) ? m : sum + m;
    }
    uint64_t denom = P.ref_count;
    if (P.ref_count > 1)
        denom -= sparse;
    if (denom == 0)
        sum = -1.0;
    scores[p] = sum < 0 ? -1 : (int32_t)floor(100.0 * sum / (double)denom + 0.5);
}



// chunk_max_kernel for 64-bit maxima (num << 32 | den, den 0 = infinity,
// ~0 = no qualifying target) of the MIXED comparison mode.
__global__ void chunk_max_wide_kernel(const CmpPair* __restrict__ pairs, uint32_t npairs, uint64_t ninst,
                                      const uint64_t* __restrict__ partial, uint64_t* __restrict__ imax)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= ninst)
        return;
    uint32_t lo = 0, hi = npairs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pairs[mid].inst_begin <= t)
            lo = mid;
        else
            hi = mid;
    }
    const CmpPair P = pairs[lo];
    const uint64_t* v = partial + P.partial_begin + (t - P.inst_begin) * P.nch;
    uint64_t best = ~0ull;
    for (uint32_t c = 0; c < P.nch; c++) {
        const uint64_t x = v[c];
        if (x == ~0ull)
            continue;
        const uint64_t n = x >> 32, d = x & 0xFFFFFFFFull, bn = best >> 32, bd = best & 0xFFFFFFFFull;
        if (best == ~0ull || (bd != 0 && (d == 0 || n * bd > bn * d)))
            best = x;
    }
    imax[t] = best;
}

// pair_sum_kernel over 64-bit maxima: each maximum is the double quotient
// num / den (+infinity for den 0), summed in filter order; a sum that ends
// infinite is flagged SDBFX_SCORE_NONFINITE (boost::math::round raises on it
// in the reference).  Scores above 100 are possible here (fractions above 1).
__global__ void pair_sum_wide_kernel(const CmpPair* __restrict__ pairs, uint32_t npairs,
                                     const uint16_t* __restrict__ ecfA, const uint16_t* __restrict__ ecfB,
                                     const uint64_t* __restrict__ imax, int32_t* __restrict__ scores)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs)
        return;
    const CmpPair P = pairs[p];
    const uint16_t* __restrict__ ecf = P.ref_set ? ecfB : ecfA;
    double sum = -1.0;
    uint32_t sparse = 0;
    for (uint32_t i = 0; i < P.ref_count; i++) {
        const uint32_t s1 = ecf[P.ref_first + i];
        double m;
        if (s1 < SDBFX_MIN_ELEM) {
            m = 0.0;
            sparse++;
        } else {
            const uint64_t x = imax[P.inst_begin + i];
            const uint32_t d = (uint32_t)x;
            m = x == ~0ull ? -1.0 : d == 0 ? __builtin_huge_val() : (double)(uint32_t)(x >> 32) / (double)d;
        }
        sum = (sum < 0) ? m : sum + m;
    }
    uint64_t denom = P.ref_count;
    if (P.ref_count > 1)
        denom -= sparse;
    if (denom == 0)
        sum = -1.0;
    const double v = 100.0 * sum / (double)denom;
    scores[p] = sum < 0 ? -1 : v == __builtin_huge_val() ? SDBFX_SCORE_NONFINITE : (int32_t)floor(v + 0.5);
}

// ------------------------------------------------------------------ text form

// The text form of a set (sdbf::to_string, sdbf/sdbf_class.cc:294-334, for
// each digest in order).  text_header_kernel places each digest's header
// (formatted on the host) and its closing newline; text_b64_kernel writes
// the bodies, one thread per base64 quantum (3 input bytes -> 4 characters,
// RFC 4648 alphabet with '=' padding, as base64/modp_b64.cc:79-118).  A
// stream-mode body is the base64 of all its filters as one run; a block-mode
// body is, per filter, ':', its element count as two lower-case hex digits,
// ':' and the 344-character base64 of the filter (86 quanta, the last one of
// a single byte), so 87 quanta per filter with the first one the prefix.
__constant__ char c_b64[65] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
__constant__ char c_hex[17] = "0123456789abcdef";

__global__ void text_header_kernel(const TextDigest* __restrict__ td, const char* __restrict__ hdr,
                                   uint64_t text_len, char* __restrict__ text)
{
    const TextDigest D = td[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < D.hdr_len; i += blockDim.x)
        text[D.text_off + i] = hdr[D.hdr_off + i];
    if (threadIdx.x == 0) {
        const uint64_t end = blockIdx.x + 1 < gridDim.x ? td[blockIdx.x + 1].text_off : text_len;
        text[end - 1] = '\n';
    }
}

__global__ void text_b64_kernel(const uint8_t* __restrict__ filt, const uint16_t* __restrict__ ecf,
                                const TextDigest* __restrict__ td, uint32_t nd, uint64_t nq,
                                char* __restrict__ text)
{
    const uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (td[mid].q_begin <= q)
            lo = mid;
        else
            hi = mid;
    }
    const TextDigest D = td[lo];
    const uint64_t lq = q - D.q_begin;
    const uint8_t* src = filt + D.filt_first * SDBFX_BF;
    uint64_t ib;
    uint32_t nb;
    char* dst;
    if (D.dd == 0) {
        const uint64_t total = (uint64_t)D.bf_count * SDBFX_BF;
        ib = 3 * lq;
        nb = (uint32_t)min((uint64_t)3, total - ib);
        dst = text + D.body_off + 4 * lq;
    } else {
        const uint64_t f = lq / 87;
        const uint32_t r = (uint32_t)(lq % 87);
        char* fd = text + D.body_off + f * 348;
        if (r == 0) {
            const uint32_t ec = ecf[D.filt_first + f];
            fd[0] = ':';
            fd[1] = c_hex[(ec >> 4) & 15];
            fd[2] = c_hex[ec & 15];
            fd[3] = ':';
            return;
        }
        ib = f * SDBFX_BF + 3 * (r - 1);
        nb = min(3u, (uint32_t)SDBFX_BF - 3 * (r - 1));
        dst = fd + 4 + 4 * (r - 1);
    }
    const uint32_t v = (uint32_t)src[ib] << 16 | (nb > 1 ? (uint32_t)src[ib + 1] << 8 : 0u) |
                       (nb > 2 ? (uint32_t)src[ib + 2] : 0u);
    dst[0] = c_b64[v >> 18 & 63];
    dst[1] = c_b64[v >> 12 & 63];
    dst[2] = nb > 1 ? c_b64[v >> 6 & 63] : '=';
    dst[3] = nb > 2 ? c_b64[v & 63] : '=';
}

// Parsing the text form (sdbf::sdbf(const std::string&), sdbf/sdbf_class.cc:
// 144-215): the host has read each line's header fields; text_decode_kernel
// decodes the bodies, one thread per base64 quantum, straight into the set
// (block mode: quantum 0 of each filter is its ':xx:' element count), and
// raises *err on a character outside the alphabet or a misplaced '='.
// text_finish_kernel then fills each new filter's Hamming weight
// (sdbf::compute_hamming, sdbf_class.cc:389-401) and, in stream mode, its
// element count (max_elem, or last_count for the last filter).
__device__ __forceinline__ int b64_value(uint32_t c)
{
    return c >= 'A' && c <= 'Z' ? (int)(c - 'A')
           : c >= 'a' && c <= 'z' ? (int)(c - 'a' + 26)
           : c >= '0' && c <= '9' ? (int)(c - '0' + 52)
           : c == '+' ? 62 : c == '/' ? 63 : -1;
}

__device__ __forceinline__ int hex_value(uint32_t c)
{
    return c >= '0' && c <= '9' ? (int)(c - '0') : c >= 'a' && c <= 'f' ? (int)(c - 'a' + 10)
           : c >= 'A' && c <= 'F' ? (int)(c - 'A' + 10) : -1;
}

__device__ __forceinline__ uint32_t find_digest(const TextParse* __restrict__ tp, uint32_t nd, uint64_t q)
{
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tp[mid].q_begin <= q)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ void text_decode_kernel(const char* __restrict__ text, const TextParse* __restrict__ tp, uint32_t nd,
                                   uint64_t nq, uint8_t* __restrict__ filt, uint16_t* __restrict__ ecf,
                                   uint32_t* __restrict__ err)
{
    const uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    const TextParse D = tp[find_digest(tp, nd, q)];
    const uint64_t lq = q - D.q_begin;
    // the quantum's 4 characters (block mode: quantum 0 of each filter is its
    // ':xx:' prefix) at body_off + 4 lq, read as the one or two aligned words
    // that hold them (the text buffer extends past the last line)
    const uint64_t A = D.body_off + 4 * lq;
    const uint32_t sh = (uint32_t)(A & 3);
    const uint32_t* wp = (const uint32_t*)(text + (A - sh));
    const uint32_t cw = sh ? (wp[0] >> (8 * sh) | wp[1] << (32 - 8 * sh)) : wp[0];
    const uint32_t ch0 = cw & 0xFFu, ch1 = cw >> 8 & 0xFFu, ch2 = cw >> 16 & 0xFFu, ch3 = cw >> 24;
    uint64_t ob;   // output byte within the digest
    uint32_t nb;
    if (D.dd == 0) {
        const uint64_t total = (uint64_t)D.bf_count * SDBFX_BF;
        ob = 3 * lq;
        nb = (uint32_t)min((uint64_t)3, total - ob);
    } else {
        const uint64_t f = lq / 87;
        const uint32_t r = (uint32_t)(lq % 87);
        if (r == 0) {
            const int h1 = hex_value(ch1), h0 = hex_value(ch2);
            if (ch0 != ':' || ch3 != ':' || h1 < 0 || h0 < 0)
                atomicOr(err, 1u);
            ecf[D.filt_first + f] = (uint16_t)((h1 & 15) << 4 | (h0 & 15));
            return;
        }
        ob = f * SDBFX_BF + 3 * (r - 1);
        nb = min(3u, (uint32_t)SDBFX_BF - 3 * (r - 1));
    }
    const int v0 = b64_value(ch0), v1 = b64_value(ch1);
    const int v2 = nb > 1 ? b64_value(ch2) : (ch2 == '=' ? 0 : -1);
    const int v3 = nb > 2 ? b64_value(ch3) : (ch3 == '=' ? 0 : -1);
    if ((v0 | v1 | v2 | v3) < 0)
        atomicOr(err, 1u);
    const uint32_t v = (uint32_t)(v0 & 63) << 18 | (uint32_t)(v1 & 63) << 12 | (uint32_t)(v2 & 63) << 6 |
                       (uint32_t)(v3 & 63);
    uint8_t* dst = filt + D.filt_first * SDBFX_BF + ob;
    dst[0] = (uint8_t)(v >> 16);
    if (nb > 1)
        dst[1] = (uint8_t)(v >> 8);
    if (nb > 2)
        dst[2] = (uint8_t)v;
}

__global__ void text_finish_kernel(const TextParse* __restrict__ tp, uint32_t nd, uint64_t f0, uint64_t nf,
                                   const uint8_t* __restrict__ filt, uint16_t* __restrict__ ecf,
                                   uint16_t* __restrict__ hamf)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= nf)
        return;
    const uint64_t fi = f0 + t;
    uint32_t lo = 0, hi = nd;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tp[mid].filt_first <= fi)
            lo = mid;
        else
            hi = mid;
    }
    const TextParse D = tp[lo];
    const uint4* w = (const uint4*)(filt + fi * SDBFX_BF);
    uint32_t h = 0;
#pragma unroll
    for (int k = 0; k < SDBFX_BF / 16; k++) {
        const uint4 x = w[k];
        h += __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
    }
    hamf[fi] = (uint16_t)h;
    if (D.dd == 0)
        ecf[fi] = (uint16_t)(fi + 1 < D.filt_first + D.bf_count ? D.max_elem : D.last_count);
}


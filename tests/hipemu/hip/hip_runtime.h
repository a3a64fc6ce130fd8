// CPU emulation of the subset of HIP used by sdhash_amd/csrc, so that the
// unmodified kernels can run on the host under AddressSanitizer: device
// memory is host memory, each workgroup runs as blockDim.x cooperative
// fibers (x86-64 only), and wave operations (shuffles, ballots, readlane) exchange values
// through per-wave slots between two wave barriers.  Workgroups of a launch
// run one after another.  Test infrastructure only.
#ifndef HIPEMU_HIP_RUNTIME_H
#define HIPEMU_HIP_RUNTIME_H
#define HIPEMU 1

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <array>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#endif

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __constant__
#define __launch_bounds__(...)
#define __align__(n) alignas(n)

struct dim3 {
    uint32_t x, y, z;
    dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
struct uint4 {
    uint32_t x, y, z, w;
};
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }

// Minimal x86-64 context switch: saves the callee-saved registers on the
// current stack, stores its pointer in *save, and resumes the stack `next`.
extern "C" void hipemu_switch(void** save, void* next);
asm(R"(
    .text
    .globl hipemu_switch
    .type hipemu_switch,@function
hipemu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size hipemu_switch, .-hipemu_switch
)");

namespace hipemu {

// Each workgroup runs as blockDim.x cooperative fibers on one host thread.
// A fiber gives up the CPU at every barrier and wave operation; the scheduler
// resumes only fibers whose barrier has opened, so the lanes of a wave stay
// in lock step as far as the kernels rely on it.
constexpr int kWave = 64;
constexpr size_t kStack = 256 * 1024;

struct Fiber {
    void* sp = nullptr;
    std::vector<char> stack;
    bool done = false;
    int wait = 0;          // 0 runnable, 1 block barrier, 2 wave barrier
    uint64_t gen = 0;      // generation it waits to see pass
};

struct Block {
    uint32_t nthreads = 0;
    std::vector<Fiber> fib;
    void* sched_sp = nullptr;
    uint32_t cur = 0;
    uint32_t bar_count = 0;
    uint64_t bar_gen = 0;
    std::vector<uint32_t> wcount;
    std::vector<uint64_t> wgen;
    std::vector<uint64_t> slot;
    std::vector<std::array<int32_t, 8>> mslot;   // matrix-core operands of each lane
    int orv = 0;
};

inline thread_local dim3 tl_threadIdx, tl_blockIdx, tl_blockDim, tl_gridDim;
inline thread_local Block* tl_block = nullptr;
inline thread_local std::function<void()>* tl_body = nullptr;
inline thread_local const void* tl_sched_bottom = nullptr;
inline thread_local size_t tl_sched_size = 0;

inline void yield()
{
    Block* b = tl_block;
    Fiber& f = b->fib[b->cur];
#if defined(__SANITIZE_ADDRESS__)
    void* fs = nullptr;
    __sanitizer_start_switch_fiber(&fs, tl_sched_bottom, tl_sched_size);
    hipemu_switch(&f.sp, b->sched_sp);
    __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
#else
    hipemu_switch(&f.sp, b->sched_sp);
#endif
}

inline int lane() { return (int)(tl_threadIdx.x % kWave); }
inline int wave() { return (int)(tl_threadIdx.x / kWave); }
inline int wave_size()
{
    const int w = wave();
    const int n = (int)tl_block->nthreads - w * kWave;
    return n < kWave ? n : kWave;
}

inline void block_sync()
{
    Block* b = tl_block;
    const uint64_t g = b->bar_gen;
    if (++b->bar_count == b->nthreads) {
        b->bar_count = 0;
        b->bar_gen++;
        return;
    }
    Fiber& f = b->fib[b->cur];
    f.wait = 1;
    f.gen = g;
    yield();
}

inline void wave_sync()
{
    Block* b = tl_block;
    const int w = wave();
    const uint64_t g = b->wgen[w];
    if (++b->wcount[w] == (uint32_t)wave_size()) {
        b->wcount[w] = 0;
        b->wgen[w]++;
        return;
    }
    Fiber& f = b->fib[b->cur];
    f.wait = 2;
    f.gen = g;
    yield();
}

// every lane of the wave publishes v; returns the value of lane src
inline uint64_t exchange(uint64_t v, int src)
{
    Block* b = tl_block;
    const int base = wave() * kWave;
    b->slot[tl_threadIdx.x] = v;
    wave_sync();
    const uint64_t r = (src >= 0 && src < wave_size()) ? b->slot[base + src] : v;
    wave_sync();
    return r;
}

inline uint64_t ballot(bool p)
{
    Block* b = tl_block;
    const int base = wave() * kWave;
    b->slot[tl_threadIdx.x] = p ? 1 : 0;
    wave_sync();
    uint64_t m = 0;
    for (int l = 0; l < wave_size(); l++)
        if (b->slot[base + l])
            m |= 1ull << l;
    wave_sync();
    return m;
}

[[noreturn]] inline void fiber_entry()
{
#if defined(__SANITIZE_ADDRESS__)
    __sanitizer_finish_switch_fiber(nullptr, &tl_sched_bottom, &tl_sched_size);
#endif
    (*tl_body)();
    Block* b = tl_block;
    Fiber& f = b->fib[b->cur];
    f.done = true;
#if defined(__SANITIZE_ADDRESS__)
    __sanitizer_start_switch_fiber(nullptr, tl_sched_bottom, tl_sched_size);
#endif
    void* dummy;
    hipemu_switch(&dummy, b->sched_sp);
    __builtin_unreachable();
}

// One kernel at a time in the process: __shared__ arrays are function
// statics here, so launches from several host threads are serialized.
inline std::mutex& launch_mutex()
{
    static std::mutex m;
    return m;
}

template <typename F>
void launch(const dim3& grid, const dim3& block, F&& body)
{
    std::lock_guard<std::mutex> serialize(launch_mutex());
    const uint32_t nt = block.x;
    std::function<void()> fn = body;
    for (uint32_t bx = 0; bx < grid.x; bx++) {
        Block b;
        b.nthreads = nt;
        b.fib.resize(nt);
        const uint32_t nw = (nt + kWave - 1) / kWave;
        b.wcount.assign(nw, 0);
        b.wgen.assign(nw, 0);
        b.slot.assign(nt, 0);
        b.mslot.assign(nt, {});
        tl_block = &b;
        tl_body = &fn;
        tl_blockIdx = dim3(bx);
        tl_blockDim = block;
        tl_gridDim = grid;
        for (uint32_t t = 0; t < nt; t++) {
            Fiber& f = b.fib[t];
            f.stack.resize(kStack);
            uintptr_t top = ((uintptr_t)f.stack.data() + kStack) & ~(uintptr_t)15;
            void** sp = (void**)top;
            *--sp = nullptr;                   // fake return address of the entry
            *--sp = (void*)&fiber_entry;       // popped by `ret`
            for (int r = 0; r < 6; r++)
                *--sp = nullptr;               // rbp rbx r12 r13 r14 r15
            f.sp = sp;
        }
        uint32_t live = nt;
        while (live) {
            bool ran = false;
            for (uint32_t t = 0; t < nt; t++) {
                Fiber& f = b.fib[t];
                if (f.done)
                    continue;
                if (f.wait == 1 && b.bar_gen == f.gen)
                    continue;
                if (f.wait == 2 && b.wgen[t / kWave] == f.gen)
                    continue;
                f.wait = 0;
                b.cur = t;
                tl_threadIdx = dim3(t);
                ran = true;
#if defined(__SANITIZE_ADDRESS__)
                void* fs = nullptr;
                __sanitizer_start_switch_fiber(&fs, f.stack.data(), kStack);
                hipemu_switch(&b.sched_sp, f.sp);
                __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
#else
                hipemu_switch(&b.sched_sp, f.sp);
#endif
                if (f.done)
                    live--;
            }
            if (!ran) {
                fprintf(stderr, "hipemu: deadlock in workgroup %u (fibers wait on barriers that cannot open)\n", bx);
                abort();
            }
        }
        tl_block = nullptr;
    }
}

}  // namespace hipemu

#define threadIdx (hipemu::tl_threadIdx)
#define blockIdx (hipemu::tl_blockIdx)
#define blockDim (hipemu::tl_blockDim)
#define gridDim (hipemu::tl_gridDim)

inline void __syncthreads() { hipemu::block_sync(); }
inline int __syncthreads_or(int p)
{
    hipemu::Block* b = hipemu::tl_block;
    hipemu::block_sync();
    if (threadIdx.x == 0)
        b->orv = 0;
    hipemu::block_sync();
    if (p)
        b->orv = 1;
    hipemu::block_sync();
    const int r = b->orv;
    hipemu::block_sync();
    return r;
}
inline void __threadfence() {}
inline void __threadfence_block() {}
inline void __builtin_amdgcn_wave_barrier() { hipemu::wave_sync(); }
inline void __builtin_amdgcn_sched_barrier(int) {}
#define __builtin_amdgcn_fence(order, scope) ((void)0)

template <typename T>
inline T __shfl_down(T v, unsigned off)
{
    return (T)hipemu::exchange((uint64_t)v, hipemu::lane() + (int)off);
}
template <typename T>
inline T __shfl_up(T v, unsigned off)
{
    return (T)hipemu::exchange((uint64_t)v, hipemu::lane() - (int)off);
}
template <typename T>
inline T __shfl(T v, int src)
{
    return (T)hipemu::exchange((uint64_t)v, src);
}
inline unsigned long long __ballot(int p) { return hipemu::ballot(p != 0); }

// DPP moves used by the kernels: row_shl:k (0x101-0x10F), row_shr:k
// (0x111-0x11F), wave_shl:1 (0x130), wave_shr:1 (0x138); full row/bank masks.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl)
{
    (void)row_mask;
    (void)bank_mask;
    const int l = hipemu::lane();
    int from = -1;
    if (ctrl >= 0x101 && ctrl <= 0x10F) {
        const int k = ctrl - 0x100;
        if ((l & 15) + k < 16)
            from = l + k;
    } else if (ctrl >= 0x111 && ctrl <= 0x11F) {
        const int k = ctrl - 0x110;
        if ((l & 15) - k >= 0)
            from = l - k;
    } else if (ctrl == 0x130) {
        from = l + 1 < hipemu::kWave ? l + 1 : -1;
    } else if (ctrl == 0x138) {
        from = l - 1;
    } else {
        fprintf(stderr, "hipemu: unsupported dpp control %#x\n", ctrl);
        abort();
    }
    const uint64_t v = hipemu::exchange((uint32_t)src, from < 0 ? 1000 : from);
    if (from < 0 || from >= hipemu::wave_size())
        return bound_ctrl ? 0 : old;
    return (int)(uint32_t)v;
}
template <typename T>
inline T __shfl_xor(T v, int mask)
{
    return (T)hipemu::exchange((uint64_t)v, hipemu::lane() ^ mask);
}

// v_mfma_i32_32x32x32_i8: A lane l = row l%32, 16 int8 of K-half l/32; B lane
// l = column l%32, same K assignment; D lane l register i = row
// 8*(i/4) + 4*(l/32) + i%4, column l%32.
typedef int hipemu_v4i __attribute__((vector_size(16)));
typedef int hipemu_v16i __attribute__((vector_size(64)));
inline hipemu_v16i __builtin_amdgcn_mfma_i32_32x32x32_i8(hipemu_v4i a, hipemu_v4i b, hipemu_v16i c, int, int, int)
{
    hipemu::Block* bk = hipemu::tl_block;
    const int base = hipemu::wave() * hipemu::kWave, l = hipemu::lane();
    auto& me = bk->mslot[threadIdx.x];
    for (int k = 0; k < 4; k++) {
        me[k] = a[k];
        me[4 + k] = b[k];
    }
    hipemu::wave_sync();
    hipemu_v16i d;
    const int n = l % 32;
    for (int i = 0; i < 16; i++) {
        const int m = 8 * (i / 4) + 4 * (l / 32) + i % 4;
        int32_t s = c[i];
        for (int hh = 0; hh < 2; hh++) {
            const int8_t* pa = (const int8_t*)bk->mslot[base + 32 * hh + m].data();
            const int8_t* pb = (const int8_t*)(bk->mslot[base + 32 * hh + n].data() + 4);
            for (int j = 0; j < 16; j++)
                s += (int32_t)pa[j] * (int32_t)pb[j];
        }
        d[i] = s;
    }
    hipemu::wave_sync();
    return d;
}

// v_mfma_scale_f32_32x32x64_f8f6f4 with FP4 (E2M1) A and B (cbsz = blgp = 4):
// A lane l = row l%32, 32 nibbles (registers 0-3, low nibble first) of
// K-half l/32; B lane l = column l%32, same K assignment; each element times
// its block scale 2^(scale - 127); D lane l register i = row
// 8*(i/4) + 4*(l/32) + i%4, column l%32.
typedef int hipemu_v8i __attribute__((vector_size(32)));
typedef float hipemu_v16f __attribute__((vector_size(64)));
inline float hipemu_fp4(uint32_t nib)
{
    static const float mag[8] = {0.f, 0.5f, 1.f, 1.5f, 2.f, 3.f, 4.f, 6.f};
    return (nib & 8) ? -mag[nib & 7] : mag[nib & 7];
}
inline hipemu_v16f __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(hipemu_v8i a, hipemu_v8i b, hipemu_v16f c,
                                                                   int cbsz, int blgp, int, int sa, int, int sb)
{
    if (cbsz != 4 || blgp != 4)
        abort();   // only the FP4 form is emulated
    hipemu::Block* bk = hipemu::tl_block;
    const int base = hipemu::wave() * hipemu::kWave, l = hipemu::lane();
    auto& me = bk->mslot[threadIdx.x];
    for (int k = 0; k < 4; k++) {
        me[k] = a[k];
        me[4 + k] = b[k];
    }
    hipemu::wave_sync();
    const float s = std::ldexp(1.0f, sa - 127) * std::ldexp(1.0f, sb - 127);
    hipemu_v16f d;
    const int n = l % 32;
    for (int i = 0; i < 16; i++) {
        const int m = 8 * (i / 4) + 4 * (l / 32) + i % 4;
        float acc = c[i];
        for (int hh = 0; hh < 2; hh++) {
            const int32_t* pa = bk->mslot[base + 32 * hh + m].data();
            const int32_t* pb = bk->mslot[base + 32 * hh + n].data() + 4;
            for (int j = 0; j < 32; j++)
                acc += hipemu_fp4(((uint32_t)pa[j / 8] >> (4 * (j % 8))) & 15u) *
                       hipemu_fp4(((uint32_t)pb[j / 8] >> (4 * (j % 8))) & 15u) * s;
        }
        d[i] = acc;
    }
    hipemu::wave_sync();
    return d;
}

inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t m, uint32_t base)
{
    const int l = hipemu::lane();
    const uint32_t below = l >= 32 ? 0xFFFFFFFFu : ((1u << l) - 1u);
    return base + (uint32_t)__builtin_popcount(m & below);
}
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t m, uint32_t base)
{
    const int l = hipemu::lane();
    const uint32_t below = l < 32 ? 0u : ((1u << (l - 32)) - 1u);
    return base + (uint32_t)__builtin_popcount(m & below);
}

inline uint32_t __lane_id() { return (uint32_t)hipemu::lane(); }
inline int __builtin_amdgcn_readlane(int v, int l) { return (int)hipemu::exchange((uint32_t)v, l); }
inline int __builtin_amdgcn_readfirstlane(int v) { return (int)hipemu::exchange((uint32_t)v, 0); }
inline uint32_t __builtin_amdgcn_alignbit(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (s & 31));
}
inline uint32_t __builtin_amdgcn_alignbyte(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (8 * (s & 3)));
}
inline int __popc(uint32_t x) { return __builtin_popcount(x); }
inline int __clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
inline uint64_t __builtin_amdgcn_s_memrealtime() { return 0; }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v)
{
    const unsigned long long o = *p;
    *p = o + v;
    return o;
}

inline uint32_t atomicAdd(uint32_t* p, uint32_t v) { const uint32_t o = *p; *p = o + v; return o; }
inline uint32_t atomicSub(uint32_t* p, uint32_t v) { const uint32_t o = *p; *p = o - v; return o; }
inline uint32_t atomicOr(uint32_t* p, uint32_t v) { const uint32_t o = *p; *p = o | v; return o; }
inline uint32_t atomicAnd(uint32_t* p, uint32_t v) { const uint32_t o = *p; *p = o & v; return o; }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v)
{
    const unsigned long long o = *p;
    if (v > o)
        *p = v;
    return o;
}
inline uint32_t atomicMin(uint32_t* p, uint32_t v)
{
    const uint32_t o = *p;
    *p = v < o ? v : o;
    return o;
}
inline uint32_t atomicCAS(uint32_t* p, uint32_t cmp, uint32_t v)
{
    const uint32_t o = *p;
    if (o == cmp)
        *p = v;
    return o;
}
#define __HIP_MEMORY_SCOPE_AGENT 3
inline uint32_t __hip_atomic_load(const uint32_t* p, int, int) { return *p; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }

inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
inline uint64_t min(uint64_t a, uint64_t b) { return a < b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }

// ------------------------------------------------------------- runtime API

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
#define hipStreamNonBlocking 1

inline const char* hipGetErrorString(hipError_t e) { return e ? "emulated HIP error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipMalloc(void** p, size_t n)
{
    *p = malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T>
inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t)
{
    memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemset(void* d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = (void*)new int(0); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete (int*)e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
// launches run to completion before they return, so every event has happened
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
#define hipEventDisableTiming 0x2
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
#define hipHostMallocDefault 0u
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) { *p = malloc(n); return *p ? hipSuccess : hipErrorOutOfMemory; }
inline hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
// a __device__ variable is host memory here: its address is the device address
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }

// graphs: launches on a capturing stream are recorded (arguments copied) and
// run by hipGraphLaunch
struct hipemu_graph {
    std::vector<std::function<void()>> ops;
};
typedef hipemu_graph* hipGraph_t;
typedef hipemu_graph* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
inline hipemu_graph*& hipemu_capturing()
{
    static thread_local hipemu_graph* g = nullptr;
    return g;
}
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
    hipemu_capturing() = new hipemu_graph;
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g)
{
    *g = hipemu_capturing();
    hipemu_capturing() = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, void*, void*, size_t)
{
    *x = new hipemu_graph(*g);
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t)
{
    for (auto& op : x->ops)
        op();
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t x) { delete x; return hipSuccess; }

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)                          \
    do {                                                                                     \
        if (getenv("HIPEMU_TRACE"))                                                          \
            fprintf(stderr, "[hipemu] %s grid %u block %u\n", #kernel, (unsigned)dim3(grid).x, \
                    (unsigned)dim3(block).x);                                                \
        if (hipemu_capturing()) {                                                            \
            const dim3 g_ = dim3(grid), b_ = dim3(block);                                    \
            hipemu_capturing()->ops.push_back(                                               \
                [=]() { hipemu::launch(g_, b_, [&]() { kernel(__VA_ARGS__); }); });          \
        } else {                                                                             \
            hipemu::launch((grid), (block), [&]() { kernel(__VA_ARGS__); });                 \
        }                                                                                    \
    } while (0)

#endif

// Device-side generator of synthetic segment-mix input for benchmarks
// (libsdbfx_synth.so; not part of the engine).  Each 4 KiB segment is, by a
// hash of (seed, segment index): 40 % uniform random bytes, 40 % Zipf(s=1.1)
// over a 64-symbol alphabet, 10 % one 16-byte motif repeated, 10 % zeros --
// the mix of sdhash_amd/synth.py, from a counter-based hash instead of
// numpy's generator, so its bytes differ from the host generator's.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    // splitmix64 finaliser
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

struct ZipfTable {
    uint32_t cdf[64];   // cumulative probability of symbols 0..k, scaled to 2^32
};

__device__ __forceinline__ uint32_t zipf_symbol(const ZipfTable& z, uint32_t u)
{
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1)
        if (lo + step - 1 < 63 && z.cdf[lo + step - 1] <= u)
            lo += step;
    return lo;
}

// bytes [offset, offset + n) of stream `seed`, written to out[0..n); 16 bytes
// per thread step; n a multiple of 16, offset a multiple of 16
__global__ void synth_fill_kernel(uint8_t* __restrict__ out, uint64_t n, uint64_t offset, uint64_t seed, ZipfTable z)
{
    for (uint64_t i = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) * 16; i < n;
         i += (uint64_t)gridDim.x * blockDim.x * 16) {
        const uint64_t pos = offset + i;
        const uint64_t seg = pos >> 12;
        const uint64_t hs = mix64(seed * 0x9e3779b97f4a7c15ull + seg);
        const uint32_t kind = (uint32_t)(hs % 10);
        uint32_t w[4];
        if (kind < 4) {
            const uint64_t a = mix64(hs ^ (pos >> 4) * 0x2545f4914f6cdd1dull);
            const uint64_t b = mix64(a + 0x632be59bd9b4e019ull);
            w[0] = (uint32_t)a;
            w[1] = (uint32_t)(a >> 32);
            w[2] = (uint32_t)b;
            w[3] = (uint32_t)(b >> 32);
        } else if (kind < 8) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t r = mix64(hs ^ ((pos >> 2) + k) * 0x2545f4914f6cdd1dull);
                uint32_t v = 0;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    v |= (0x20u + zipf_symbol(z, (uint32_t)(r >> (16 * j)) << 16 | (uint32_t)(r >> 48))) << (8 * j);
                w[k] = v;
            }
        } else if (kind == 8) {
            const uint64_t m0 = mix64(hs + 1), m1 = mix64(hs + 2);
            w[0] = (uint32_t)m0;
            w[1] = (uint32_t)(m0 >> 32);
            w[2] = (uint32_t)m1;
            w[3] = (uint32_t)(m1 >> 32);
        } else {
            w[0] = w[1] = w[2] = w[3] = 0;
        }
        *(uint4*)(out + i) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Planted chunks: chunk c copies len[c] bytes from src + s_off[c] to
// dst + d_off[c], one workgroup per chunk, 16 bytes per thread step (offsets
// and lengths multiples of 16; the source is not written by any chunk).
__global__ void synth_plant_kernel(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src,
                                   const uint64_t* __restrict__ d_off, const uint64_t* __restrict__ s_off,
                                   const uint32_t* __restrict__ len)
{
    const uint64_t d = d_off[blockIdx.x], sidx = s_off[blockIdx.x];
    const uint32_t n = len[blockIdx.x];
    for (uint32_t i = threadIdx.x * 16; i < n; i += blockDim.x * 16)
        *(uint4*)(dst + d + i) = *(const uint4*)(src + sidx + i);
}

extern "C" {

// Copies n chunks from the device buffer src into the device buffer dst
// (host arrays of offsets and lengths, all multiples of 16).  Returns 0 or -1.
int sdbfx_synth_plant(void* dst, const void* src, uint32_t n, const uint64_t* d_off, const uint64_t* s_off,
                      const uint32_t* len)
{
    if (!dst || !src || !n)
        return n ? -1 : 0;
    for (uint32_t c = 0; c < n; c++)
        if ((d_off[c] | s_off[c] | len[c]) & 15)
            return -1;
    uint64_t* dd = nullptr;
    uint64_t* ds = nullptr;
    uint32_t* dl = nullptr;
    int rc = -1;
    if (hipMalloc(&dd, n * 8ull) == hipSuccess && hipMalloc(&ds, n * 8ull) == hipSuccess &&
        hipMalloc(&dl, n * 4ull) == hipSuccess && hipMemcpy(dd, d_off, n * 8ull, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(ds, s_off, n * 8ull, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(dl, len, n * 4ull, hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(synth_plant_kernel, dim3(n), dim3(256), 0, 0, (uint8_t*)dst, (const uint8_t*)src, dd, ds,
                           dl);
        if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess)
            rc = 0;
    }
    (void)hipFree(dd);
    (void)hipFree(ds);
    (void)hipFree(dl);
    return rc;
}

// Fills device memory out[0..bytes) with the stream `seed` from byte
// `offset` on; bytes and offset multiples of 16.  Returns 0 or -1.
int sdbfx_synth_fill(void* out, uint64_t bytes, uint64_t offset, uint64_t seed)
{
    if (!out || (bytes & 15) || (offset & 15))
        return -1;
    ZipfTable z;
    double norm = 0, c = 0;
    for (int k = 1; k <= 64; k++)
        norm += 1.0 / std::pow((double)k, 1.1);
    for (int k = 1; k <= 64; k++) {
        c += 1.0 / std::pow((double)k, 1.1) / norm;
        const double v = c * 4294967296.0;
        z.cdf[k - 1] = v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
    }
    const uint64_t threads = (bytes / 16 + 255) / 256;
    const uint32_t grid = (uint32_t)(threads < 65536 ? (threads ? threads : 1) : 65536);
    hipLaunchKernelGGL(synth_fill_kernel, dim3(grid), dim3(256), 0, 0, (uint8_t*)out, bytes, offset, seed, z);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return -1;
    return 0;
}

}  // extern "C"

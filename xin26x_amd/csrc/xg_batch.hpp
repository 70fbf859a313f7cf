// xg_batch.hpp -- helpers shared by the batched block kernels (forward
// transform, reconstruction): group staging of same-shape blocks into LDS,
// byte packing for int8 MFMA operands, launch sizing and device arrays.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xg_runtime.hpp"

namespace xg {

// low bytes of four int32, in order, as one dword
__device__ __forceinline__ uint32_t bytes4(int a, int b, int c, int d)
{
    const uint32_t lo = __builtin_amdgcn_perm((uint32_t)b, (uint32_t)a, 0x0c0c0400u);
    const uint32_t hi = __builtin_amdgcn_perm((uint32_t)d, (uint32_t)c, 0x0c0c0400u);
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// low bytes of the two 16-bit halves of four dwords, in order, as 8 bytes
__device__ __forceinline__ long bytes8_lo16(const uint32_t (&m)[4])
{
    return (long)(((unsigned long long)__builtin_amdgcn_perm(m[3], m[2], 0x06040200u) << 32) |
                  __builtin_amdgcn_perm(m[1], m[0], 0x06040200u));
}

__device__ __forceinline__ long load8(const int8_t *p)
{
    return *(const long *)p;      // 8-byte aligned: matrix rows are multiples of 32, k offsets of 8
}

// Chunks of a group of same-shape blocks: `cs` samples per lane and load, so
// every load is aligned (block offsets are multiples of cs); chunk c of the
// group is sample c * cs of the LDS copy, blocks back to back.  Lane b holds
// block b's offset relative to the group base.  Lanes past the group re-read
// its last chunk: every load is issued and none waits for another.
// T threads share the work (thread t issues chunks t, t + T, ...); myRel is
// read from the calling wave's lanes.
template <typename V, int N, int T = 64>
__device__ __forceinline__ void stage_group(V (&c)[N], const int16_t *base, int myRel, int lgCpb, int nch, int cs,
                                            int tid)
{
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int ch = min(tid + T * i, nch - 1);
        const int o = __shfl(myRel, ch >> lgCpb);
        c[i] = *reinterpret_cast<const V *>(base + o + (ch & ((1 << lgCpb) - 1)) * cs);
    }
}

// every slot is written (past the group: copies of its last chunk), so the
// loads stay in one straight run ahead of the stores
template <typename V, int N, int T = 64>
__device__ __forceinline__ void stage_store(const V (&c)[N], void *lds, int tid)
{
#pragma unroll
    for (int i = 0; i < N; i++) reinterpret_cast<V *>(lds)[tid + T * i] = c[i];
}

// How the WAVES waves of a workgroup share a stage of `tiles` 16-wide tiles
// with up to four 16-wide sub-tiles each (mi): whole tiles per wave when there
// are at least as many tiles as waves (or the tiles do not divide the waves),
// else the waves of one tile take every (WAVES / tiles)-th sub-tile.
struct Split {
    int t0, tstep, m0, mmask;
};

template <int WAVES>
__device__ __forceinline__ Split work_split(int tiles, int wave)
{
    Split s;
    if (tiles >= WAVES || (WAVES % tiles) != 0) {
        s.t0 = wave;
        s.tstep = WAVES;
        s.m0 = 0;
        s.mmask = 0;
    } else {
        s.t0 = wave % tiles;
        s.tstep = WAVES;                          // one tile per wave
        s.m0 = wave / tiles;
        s.mmask = WAVES / tiles - 1;              // sub-tiles m0, m0 + WAVES/tiles, ...
    }
    return s;
}

// sub-tile mi is this wave's share of its tile
__device__ __forceinline__ bool split_takes(const Split &s, int mi)
{
    return mi >= s.m0 && ((mi - s.m0) & s.mmask) == 0;
}

// workgroups of `threads` threads of `kernel` that fit on the device at once
inline int resident_slots(const void *kernel, int threads = 64)
{
    int dev = 0, cus = 256, per = 8;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        cus = prop.multiProcessorCount;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1)
        per = threads <= 64 ? 8 : 2;
    return cus * per;
}

// grows a device array to hold n elements (contents not kept)
template <typename T>
bool grow(T *&ptr, int &cap, int n, hipStream_t st)
{
    if (n <= cap) return true;
    if (ptr) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(ptr);
    }
    ptr = nullptr;
    cap = 0;
    if (!XG_HIP(hipMalloc((void **)&ptr, sizeof(T) * (size_t)n))) return false;
    cap = n;
    return true;
}

}  // namespace xg

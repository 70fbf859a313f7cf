// xg_device.hpp -- device-side building blocks shared by every gfx950 kernel:
// VVC filter tables, the separable interpolation arithmetic, wave-level
// shuffles and Hadamard butterflies.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xg {

constexpr int kPixMax = 1023;                 // 10-bit samples
// Interpolation constants for 10-bit (h266_core/include/h266_inter_pred_context.h:31-39):
// first pass (sum - (8192 << 2)) >> 2 into int16; second pass
// (sum + 512 + (8192 << 6)) >> 10 clipped to [0,1023] for PIXEL output, or
// sum >> 6 for S16 output.
constexpr int kHShift = 2;
constexpr int kHOffset = -(8192 << kHShift);
constexpr int kVShift = 10;
constexpr int kVOffset = (1 << (kVShift - 1)) + (8192 << 6);

// Filter tables live in the constant segment of every code object that
// includes this header.  [fi][phase][tap]: fi 0 default, 1 with the
// alternative half-pel filter at phase 8, 2 the 4x4-affine variant
// (h266_core/source/h266_luma_mc.c:27-70, selection rule :117-128).
static __constant__ int8_t c_luma[3][16][8] = {
    {{0, 0, 0, 64, 0, 0, 0, 0}, {0, 1, -3, 63, 4, -2, 1, 0}, {-1, 2, -5, 62, 8, -3, 1, 0},
     {-1, 3, -8, 60, 13, -4, 1, 0}, {-1, 4, -10, 58, 17, -5, 1, 0}, {-1, 4, -11, 52, 26, -8, 3, -1},
     {-1, 3, -9, 47, 31, -10, 4, -1}, {-1, 4, -11, 45, 34, -10, 4, -1}, {-1, 4, -11, 40, 40, -11, 4, -1},
     {-1, 4, -10, 34, 45, -11, 4, -1}, {-1, 4, -10, 31, 47, -9, 3, -1}, {-1, 3, -8, 26, 52, -11, 4, -1},
     {0, 1, -5, 17, 58, -10, 4, -1}, {0, 1, -4, 13, 60, -8, 3, -1}, {0, 1, -3, 8, 62, -5, 2, -1},
     {0, 1, -2, 4, 63, -3, 1, 0}},
    {{0, 0, 0, 64, 0, 0, 0, 0}, {0, 1, -3, 63, 4, -2, 1, 0}, {-1, 2, -5, 62, 8, -3, 1, 0},
     {-1, 3, -8, 60, 13, -4, 1, 0}, {-1, 4, -10, 58, 17, -5, 1, 0}, {-1, 4, -11, 52, 26, -8, 3, -1},
     {-1, 3, -9, 47, 31, -10, 4, -1}, {-1, 4, -11, 45, 34, -10, 4, -1}, {0, 3, 9, 20, 20, 9, 3, 0},
     {-1, 4, -10, 34, 45, -11, 4, -1}, {-1, 4, -10, 31, 47, -9, 3, -1}, {-1, 3, -8, 26, 52, -11, 4, -1},
     {0, 1, -5, 17, 58, -10, 4, -1}, {0, 1, -4, 13, 60, -8, 3, -1}, {0, 1, -3, 8, 62, -5, 2, -1},
     {0, 1, -2, 4, 63, -3, 1, 0}},
    {{0, 0, 0, 64, 0, 0, 0, 0}, {0, 1, -3, 63, 4, -2, 1, 0}, {0, 1, -5, 62, 8, -3, 1, 0},
     {0, 2, -8, 60, 13, -4, 1, 0}, {0, 3, -10, 58, 17, -5, 1, 0}, {0, 3, -11, 52, 26, -8, 2, 0},
     {0, 2, -9, 47, 31, -10, 3, 0}, {0, 3, -11, 45, 34, -10, 3, 0}, {0, 3, -11, 40, 40, -11, 3, 0},
     {0, 3, -10, 34, 45, -11, 3, 0}, {0, 3, -10, 31, 47, -9, 2, 0}, {0, 2, -8, 26, 52, -11, 3, 0},
     {0, 1, -5, 17, 58, -10, 3, 0}, {0, 1, -4, 13, 60, -8, 2, 0}, {0, 1, -3, 8, 62, -5, 1, 0},
     {0, 1, -2, 4, 63, -3, 1, 0}}};

// h266_core/source/h266_chroma_mc.c:26-60
static __constant__ int8_t c_chroma[32][4] = {
    {0, 64, 0, 0}, {-1, 63, 2, 0}, {-2, 62, 4, 0}, {-2, 60, 7, -1}, {-2, 58, 10, -2},
    {-3, 57, 12, -2}, {-4, 56, 14, -2}, {-4, 55, 15, -2}, {-4, 54, 16, -2}, {-5, 53, 18, -2},
    {-6, 52, 20, -2}, {-6, 49, 24, -3}, {-6, 46, 28, -4}, {-5, 44, 29, -4}, {-4, 42, 30, -4},
    {-4, 39, 33, -4}, {-4, 36, 36, -4}, {-4, 33, 39, -4}, {-4, 30, 42, -4}, {-4, 29, 44, -5},
    {-4, 28, 46, -6}, {-3, 24, 49, -6}, {-2, 20, 52, -6}, {-2, 18, 53, -5}, {-2, 16, 54, -4},
    {-2, 15, 55, -4}, {-2, 14, 56, -4}, {-2, 12, 57, -3}, {-2, 10, 58, -2}, {-1, 7, 60, -2},
    {0, 4, 62, -2}, {0, 2, 63, -1}};

template <int TAPS>
__device__ __forceinline__ const int8_t *taps_for(int phase, int fi)
{
    if constexpr (TAPS == 8) return c_luma[fi][phase];
    else return c_chroma[phase];
}

__device__ __forceinline__ int clip_pix(int v) { return min(max(v, 0), kPixMax); }

// First-pass value of one row: the reference stores it as int16 (its range is
// about +/-14.3K for 10-bit input, so the narrowing never changes it).
template <int TAPS, typename P>
__device__ __forceinline__ int hpass(const P *p, const int8_t *c)
{
    int s = 0;
#pragma unroll
    for (int j = 0; j < TAPS; j++) s += (int)p[j] * (int)c[j];
    return (int)(int16_t)((s + kHOffset) >> kHShift);
}

// Second pass to PIXEL or S16.
__device__ __forceinline__ int vfinal_pix(int s) { return clip_pix((s + kVOffset) >> kVShift); }
__device__ __forceinline__ int vfinal_s16(int s) { return (int)(int16_t)(s >> 6); }

// ---------------------------------------------------------------------------
// Wave-level exchange.  lane ^ M within a wavefront of 64.
// M = 1, 2: DPP quad permutes; M = 4, 8, 16: ds_swizzle bit-mask mode
// (and_mask 0x1F, xor_mask M; stays inside each 32-lane half);
// M = 32: ds_bpermute.
// ---------------------------------------------------------------------------
template <int M>
__device__ __forceinline__ int shfl_xor(int v)
{
    if constexpr (M == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);
    else if constexpr (M == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);
    else if constexpr (M == 4) return __builtin_amdgcn_ds_swizzle(v, (4 << 10) | 0x1F);
    else if constexpr (M == 8) return __builtin_amdgcn_ds_swizzle(v, (8 << 10) | 0x1F);
    else if constexpr (M == 16) return __builtin_amdgcn_ds_swizzle(v, (16 << 10) | 0x1F);
    else return __shfl_xor(v, M, 64);
}

// Sum over aligned groups of G lanes (G power of two <= 64); every lane of the
// group receives the group total.
template <int G>
__device__ __forceinline__ int group_sum(int v)
{
    if constexpr (G >= 2) v += shfl_xor<1>(v);
    if constexpr (G >= 4) v += shfl_xor<2>(v);
    if constexpr (G >= 8) v += shfl_xor<4>(v);
    if constexpr (G >= 16) v += shfl_xor<8>(v);
    if constexpr (G >= 32) v += shfl_xor<16>(v);
    if constexpr (G >= 64) v += shfl_xor<32>(v);
    return v;
}

// One butterfly stage of a Walsh-Hadamard transform across lanes: the lane
// with bit M clear keeps a + b, the partner keeps a - b.
template <int M>
__device__ __forceinline__ int wht_lane_stage(int v, int lane)
{
    int o = shfl_xor<M>(v);
    return (lane & M) ? (o - v) : (v + o);
}

// In-register 1-D WHT of N values (natural order).
template <int N>
__device__ __forceinline__ void wht_reg(int *d)
{
#pragma unroll
    for (int len = 1; len < N; len <<= 1)
#pragma unroll
        for (int i = 0; i < N; i += len << 1)
#pragma unroll
            for (int j = i; j < i + len; j++) {
                int a = d[j], b = d[j + len];
                d[j] = a + b;
                d[j + len] = a - b;
            }
}

// SATD of one N x N tile held as one column per lane: lanes (g*N .. g*N+N-1)
// each own N vertical differences d[0..N-1].  Returns the normalised tile SATD
// (HM xCalcHADs8x8/4x4/2x2 normalisation) in every lane of the group.
template <int N>
__device__ __forceinline__ int tile_satd(int *d, int lane)
{
    wht_reg<N>(d);                       // vertical transform in registers
    int s = 0;
#pragma unroll
    for (int r = 0; r < N; r++) {        // horizontal transform across lanes
        int v = d[r];
        if constexpr (N >= 2) v = wht_lane_stage<1>(v, lane);
        if constexpr (N >= 4) v = wht_lane_stage<2>(v, lane);
        if constexpr (N >= 8) v = wht_lane_stage<4>(v, lane);
        s += abs(v);
    }
    s = group_sum<N>(s);
    if constexpr (N == 8) return (s + 2) >> 2;
    else if constexpr (N == 4) return (s + 1) >> 1;
    else return s;
}

// ---------------------------------------------------------------------------
// Packed 16-bit pairs: two samples (or two int16 values) per dword.
// ---------------------------------------------------------------------------
typedef short v2s __attribute__((ext_vector_type(2)));
// Global-memory typed pointers: loads through them stay global_load (a
// generic pointer, e.g. one passed through an empty asm, becomes flat_load).
typedef const __attribute__((address_space(1))) uint32_t gdword;
typedef const __attribute__((address_space(1))) uint16_t gword;

__device__ __forceinline__ v2s as_v2s(uint32_t v) { return __builtin_bit_cast(v2s, v); }
__device__ __forceinline__ uint32_t as_u32(v2s v) { return __builtin_bit_cast(uint32_t, v); }
// (lo & 0xFFFF) | (hi << 16) in one v_perm_b32: selector bytes 0-1 take lo's
// low bytes, 4-5 hi's.
__device__ __forceinline__ uint32_t pack16(int lo, int hi)
{
    return __builtin_amdgcn_perm((uint32_t)hi, (uint32_t)lo, 0x05040100u);
}
__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int c)
{
    return __builtin_amdgcn_sdot2(as_v2s(a), as_v2s(b), c, false);
}

// Offset-free form of the two passes, for kernels whose first-pass values are
// only intermediates; the results are the same.  u = (sum + 32) >> 2 is the
// standard first-pass value plus 8192 + 8 exactly ((sum - (8192 << 2)) >> 2 =
// (sum >> 2) - 8192, and 32 is a multiple of 4).  Every vertical filter's
// taps sum to 64, so the second pass over u already holds
// 64 * (8192 + 8) = (8192 << 6) + 512 in its sum and is sum >> 10 alone.  For
// 10-bit input u lies in [-6138, 22514] (luma; chroma inside that), so it
// still packs as int16.  Both accumulators start from inline constants (32, 0).
constexpr int kHRound = 32;
// With the vertical taps scaled by 64 (|tap| <= 4096) the second-pass sum is
// 64 times as large (|sum| < 2^28): clamped to [0, (1023 << 16) | 0xFFFF] its
// upper half is clip(sum >> 10, 0, 1023), and two such halves pack with one
// byte permute.
constexpr int kClipHi16 = (kPixMax << 16) | 0xFFFF;
__device__ __forceinline__ int clip_hi16(int s) { return min(max(s, 0), kClipHi16); }
__device__ __forceinline__ uint32_t pack_hi16(int lo, int hi)
{
    return __builtin_amdgcn_perm((uint32_t)hi, (uint32_t)lo, 0x07060302u);
}
// First product of a dot2 chain with an inline-constant start.  The clamp
// bit (saturation of the i32 sum, which these sums never reach) keeps the
// instruction in its three-operand form, so the constant needs no register.
__device__ __forceinline__ int dot2_first(uint32_t a, uint32_t b, int c)
{
    return __builtin_amdgcn_sdot2(as_v2s(a), as_v2s(b), c, true);
}
// First pass with its taps scaled by 64 as well (|tap| <= 4096, |sum| < 2^23):
// u is bits 8..23 of the scaled sum (64 * (sum + 32)) >> 8, so the u of two
// rows pack into one dword with one byte permute (bytes 1 and 2 of each) and
// no shifts.
constexpr int kHRound64 = kHRound << 6;
__device__ __forceinline__ uint32_t pack_mid16(int lo, int hi)
{
    return __builtin_amdgcn_perm((uint32_t)hi, (uint32_t)lo, 0x06050201u);
}
__device__ __forceinline__ uint32_t align16(uint32_t hi, uint32_t lo, uint32_t sh)
{
    return __builtin_amdgcn_alignbit(hi, lo, sh);
}

// Block-reduce helper for one workgroup of 256 threads.
__device__ __forceinline__ uint32_t block_sum256(uint32_t v)
{
    __shared__ uint32_t red[4];
    v = (uint32_t)group_sum<64>((int)v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    uint32_t t = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return t;
}

// Exp-Golomb se(v) length: k = v > 0 ? 2v-1 : -2v; 2*floor(log2(k+1)) + 1.
__device__ __forceinline__ uint32_t eg_bits(int v)
{
    uint32_t k = v > 0 ? (uint32_t)(2 * v - 1) : (uint32_t)(-2 * v);
    return 2u * (31u - (uint32_t)__builtin_clz(k + 1)) + 1u;
}

}  // namespace xg

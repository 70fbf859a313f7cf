// xg_mctf.hip -- motion estimation of the motion-compensated temporal filter.
//
// The reference's MCTF motion search (Xin26xMctfMotionEstFrame,
// h26x_core/include/h26x_mctf.h:27, called per reference, search level and
// block size, with the subsampling Xin26xMctfSubSample, :35, and the MV
// fields l0Mv / l1Mv / l2Mv / mv of h26x_mctf_struct.h:117-127) is closed, so
// this file is a design of its own in that role, hierarchical as the
// published temporal filter of VVC's reference software is:
//   * subsampling: each sample of a half-resolution level is the rounded mean
//     (a + b + c + d + 2) >> 2 of its 2x2 block; the level is edge-padded;
//   * a level: for every block wholly inside the picture, the MVs of the 3x3
//     lower-level blocks around (bx / 2, by / 2), scaled by the resolution
//     factor, and the zero MV are costed (rows then columns), then the integer
//     positions +/-range around the best (rows then columns), then, at the
//     finest level, 7x7 grids of quarter-pel and of 1/16-pel steps around the
//     best; the first strictly lowest SSD wins each comparison.  Fractional
//     predictions use the standard's 8-tap luma filter (h266_luma_mc.c), so
//     every SSD is of a prediction the rest of the library can reproduce.  A
//     position whose 8-tap reach [pos - 3, pos + bs + 4] leaves the
//     reference's padding is not costed.
//   * the frame: 4x level (16x16 blocks, +/-8), 2x level (16x16, +/-5, lower
//     MVs x 2), full resolution (16x16, +/-5, x 2), then 8x8 blocks at full
//     resolution (range 0, lower MVs x 1, fractional rounds).
//
// Mapping: one 64-lane workgroup per block.  Candidates are costed one by one
// from a staged (bs + 7)^2 window with the two-pass filter (exact for integer
// MVs too); the integer search stages its (2 range + bs)^2 window once and
// each lane costs raster positions, reduced by a 64-bit (SSD << 32 | index)
// minimum; each fractional round gives one position to each of 49 lanes,
// which filters its block column by column from one staged window that
// covers both rounds.
#include <algorithm>
#include <mutex>
#include <vector>

#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

constexpr int kMctfMaxRange = 8;
constexpr int kMctfMaxBs = 16;

struct MctfArgs {
    const uint16_t *cur;
    intptr_t cs;
    const uint16_t *ref;
    intptr_t rs;
    int w, h, pad;
    int bs, range, frac, factor;
    const int2 *lower;
    int lowerBw, lowerBh;
    int2 *mv;
    uint32_t *err;
    int nbx;
};

// Shared first pass of a fractional round of 8x8 blocks: 7 horizontal
// offsets x 16 row pairs x 8 columns, each dword the first-pass values of two
// consecutive window rows (low half the upper row) of one column.
constexpr int kMctfFracBs = 8;
constexpr int kMctfFracPairs = kMctfFracBs + 8;    // rows bs + 7 + up to 2 of spread, as pairs

struct MctfLds {
    uint16_t win[(2 * kMctfMaxRange + kMctfMaxBs) * (2 * kMctfMaxRange + kMctfMaxBs)];  // 32 x 32
    int16_t tmp[(kMctfMaxBs + 7) * kMctfMaxBs];
    uint16_t blk[kMctfMaxBs * kMctfMaxBs];
};

// Whether the 8-tap reach of a prediction at MV (mx, my) for the block at
// (x, y) stays inside the reference's padding.
__device__ __forceinline__ bool mctf_valid(const MctfArgs &a, int x, int y, int mx, int my)
{
    const int px = x + (mx >> 4), py = y + (my >> 4);
    return px - 3 >= -a.pad && py - 3 >= -a.pad && px + a.bs + 4 <= a.w - 1 + a.pad &&
           py + a.bs + 4 <= a.h - 1 + a.pad;
}

// A reference sample with its coordinates kept inside the padded plane
// (staging may cover cells only invalid positions would read).
__device__ __forceinline__ uint16_t ref_at(const MctfArgs &a, int x, int y)
{
    x = min(max(x, -a.pad), a.w - 1 + a.pad);
    y = min(max(y, -a.pad), a.h - 1 + a.pad);
    return a.ref[(intptr_t)y * a.rs + x];
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long k)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const unsigned long long o = (unsigned long long)__shfl_xor((long long)k, m, 64);
        k = o < k ? o : k;
    }
    return k;
}

// SSD of the block (staged in L.blk) against the prediction at MV (mx, my),
// computed by the whole wave; UINT32_MAX when the position is not costed.
__device__ uint32_t mctf_candidate(const MctfArgs &a, MctfLds &L, int x, int y, int mx, int my)
{
    if (!mctf_valid(a, x, y, mx, my)) return 0xFFFFFFFFu;
    const int lane = threadIdx.x, bs = a.bs, ww = bs + 7;
    const int px = x + (mx >> 4) - 3, py = y + (my >> 4) - 3;
    const int8_t *hf = taps_for<8>(mx & 15, 0), *vf = taps_for<8>(my & 15, 0);
    __syncthreads();
    for (int i = lane; i < ww * ww; i += 64) {
        const int r = i / ww, c = i - r * ww;
        L.win[r * 32 + c] = ref_at(a, px + c, py + r);
    }
    __syncthreads();
    for (int i = lane; i < ww * bs; i += 64) {
        const int r = i / bs, c = i - r * bs;
        L.tmp[r * kMctfMaxBs + c] = (int16_t)hpass<8>(&L.win[r * 32 + c], hf);
    }
    __syncthreads();
    uint32_t e = 0;
    for (int i = lane; i < bs * bs; i += 64) {
        const int r = i / bs, c = i - r * bs;
        int s = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) s += (int)vf[k] * (int)L.tmp[(r + k) * kMctfMaxBs + c];
        const int d = (int)L.blk[r * kMctfMaxBs + c] - vfinal_pix(s);
        e += (uint32_t)(d * d);
    }
    return (uint32_t)group_sum<64>((int)e);
}

// FRAC: the finest level's fractional rounds (a separate instantiation, so the
// integer levels keep the register budget of their own code).
// SSD of the staged block against the window at (dx, dy) on packed pairs: the
// window pairs realigned by one sample for odd dx (alignbit by 16 dx & 16; an
// even dx also reads the dword past its last pair, inside the LDS struct, and
// shifts it out), each pair difference squared and summed by one v_dot2.
template <int BS>
__device__ __forceinline__ uint32_t int_ssd(const MctfLds &L, int dx, int dy)
{
    const uint32_t *w32 = reinterpret_cast<const uint32_t *>(L.win) + (dx >> 1);
    const uint32_t *b32 = reinterpret_cast<const uint32_t *>(L.blk);
    const uint32_t sh = (uint32_t)(dx & 1) << 4;
    int e = 0;
#pragma unroll 4
    for (int r = 0; r < BS; r++) {
        const uint32_t *wr = w32 + (dy + r) * 16, *br = b32 + r * (kMctfMaxBs / 2);
#pragma unroll
        for (int j = 0; j < BS / 2; j++) {
            const v2s d = as_v2s(br[j]) - as_v2s(align16(wr[j + 1], wr[j], sh));
            e = dot2(as_u32(d), as_u32(d), e);
        }
    }
    return (uint32_t)e;
}

template <bool FRAC>
__global__ void __launch_bounds__(64) k_mctf_level(MctfArgs a)
{
    __shared__ MctfLds L;
    const int lane = threadIdx.x, bs = a.bs;
    const int bx = (int)blockIdx.x % a.nbx, by = (int)blockIdx.x / a.nbx;
    const int x = bx * bs, y = by * bs;
    for (int i = lane; i < bs * bs; i += 64) {
        const int r = i / bs, c = i - r * bs;
        L.blk[r * kMctfMaxBs + c] = a.cur[(intptr_t)(y + r) * a.cs + x + c];
    }
    __syncthreads();
    uint32_t best = 0xFFFFFFFFu;
    int bmx = 0, bmy = 0;
    // candidates: the 3x3 lower-level MVs (rows then columns), then zero
    if (a.lower) {
        int cxs[10], cys[10];
        bool has[10], allInt = true;
#pragma unroll
        for (int k = 0; k < 10; k++) {
            int cx = 0, cy = 0;
            has[k] = true;
            if (k < 9) {
                const int tx = bx / 2 + k % 3 - 1, ty = by / 2 + k / 3 - 1;
                has[k] = tx >= 0 && ty >= 0 && tx < a.lowerBw && ty < a.lowerBh;
                if (has[k]) {
                    const int2 m = a.lower[(intptr_t)ty * a.lowerBw + tx];
                    cx = m.x * a.factor;
                    cy = m.y * a.factor;
                }
            }
            cxs[k] = cx;
            cys[k] = cy;
            allInt = allInt && (!has[k] || ((cx | cy) & 15) == 0);
        }
        if (allInt) {
            // integer positions (the lower levels' MVs, scaled, always are): the prediction is the
            // reference itself, so every lane loads its samples of all ten positions at once and
            // the SSDs need no staging
            constexpr int SPL = kMctfMaxBs * kMctfMaxBs / 64;
            bool ok[10];
            uint16_t v[10][SPL];
#pragma unroll
            for (int k = 0; k < 10; k++) {
                ok[k] = has[k] && mctf_valid(a, x, y, cxs[k], cys[k]);
                const uint16_t *rp = a.ref + (intptr_t)(y + (cys[k] >> 4)) * a.rs + x + (cxs[k] >> 4);
#pragma unroll
                for (int t = 0; t < SPL; t++) {
                    const int i = lane + 64 * t, r = i / bs, c = i - r * bs;
                    v[k][t] = (ok[k] && i < bs * bs) ? rp[(intptr_t)r * a.rs + c] : (uint16_t)0;
                }
            }
#pragma unroll
            for (int k = 0; k < 10; k++) {
                int e = 0;
#pragma unroll
                for (int t = 0; t < SPL; t++) {
                    const int i = lane + 64 * t, r = i / bs, c = i - r * bs;
                    if (i < bs * bs) {
                        const int d = (int)L.blk[r * kMctfMaxBs + c] - (int)v[k][t];
                        e += d * d;
                    }
                }
                e = group_sum<64>(e);
                if (ok[k] && (uint32_t)e < best) {
                    best = (uint32_t)e;
                    bmx = cxs[k];
                    bmy = cys[k];
                }
            }
        } else {
            for (int k = 0; k < 10; k++) {
                if (!has[k]) continue;
                const uint32_t e = mctf_candidate(a, L, x, y, cxs[k], cys[k]);
                if (e < best) {
                    best = e;
                    bmx = cxs[k];
                    bmy = cys[k];
                }
            }
        }
    }
    // integer search +/-range around the best (C division: toward zero)
    {
        const int R = a.range, ix = bmx / 16, iy = bmy / 16, ww = 2 * R + bs;
        __syncthreads();
        for (int i = lane; i < ww * ww; i += 64) {
            const int r = i / ww, c = i - r * ww;
            L.win[r * 32 + c] = ref_at(a, x + ix - R + c, y + iy - R + r);
        }
        __syncthreads();
        const int n = (2 * R + 1) * (2 * R + 1);
        unsigned long long k = ~0ull;
        for (int p = lane; p < n; p += 64) {
            const int dy = p / (2 * R + 1), dx = p - dy * (2 * R + 1);
            if (!mctf_valid(a, x, y, 16 * (ix - R + dx), 16 * (iy - R + dy))) continue;
            const uint32_t e = bs == 16 ? int_ssd<16>(L, dx, dy) : int_ssd<8>(L, dx, dy);
            const unsigned long long kk = ((unsigned long long)e << 32) | (uint32_t)p;
            k = kk < k ? kk : k;
        }
        k = wave_min_u64(k);
        if (k != ~0ull && (uint32_t)(k >> 32) < best) {
            const int p = (int)(uint32_t)k, dy = p / (2 * R + 1), dx = p - dy * (2 * R + 1);
            best = (uint32_t)(k >> 32);
            bmx = 16 * (ix - R + dx);
            bmy = 16 * (iy - R + dy);
        }
    }
    // fractional rounds: 7x7 at quarter-pel steps, then 7x7 at 1/16 pel
    if constexpr (FRAC) {
        __shared__ uint32_t hp[7][kMctfFracPairs][kMctfFracBs];
        const int x0 = ((bmx - 15) >> 4) - 3, y0 = ((bmy - 15) >> 4) - 3;   // window covering both rounds
        const int ww = ((bmx + 15) >> 4) - ((bmx - 15) >> 4) + bs + 7;
        const int wh = ((bmy + 15) >> 4) - ((bmy - 15) >> 4) + bs + 7;
        __syncthreads();
        for (int i = lane; i < ww * wh; i += 64) {
            const int r = i / ww, c = i - r * ww;
            L.win[r * 32 + c] = ref_at(a, x + x0 + c, y + y0 + r);
        }
        __syncthreads();
        for (int step = 4; step >= 1; step -= 3) {
            const int ox = bmx, oy = bmy;
            unsigned long long k = ~0ull;
            if (bs == kMctfFracBs) {
                // the first pass of each of the 7 horizontal offsets, once for the 7 positions
                // that share it: rows from the highest position's first tap row (spread <= 2)
                const int rb = ((oy - 3 * step) >> 4) - 3 - y0;
                const int nr = ((oy + 3 * step) >> 4) - 3 - y0 - rb + bs + 7;
                __syncthreads();                      // earlier readers of hp are done
                for (int t = lane; t < 7 * nr * bs; t += 64) {
                    const int i = t / (nr * bs), rc = t - i * (nr * bs), r = rc / bs, c = rc - r * bs;
                    const int mx = ox + (i - 3) * step;
                    const int cx = (mx >> 4) - 3 - x0;
                    const int u = hpass<8>(&L.win[(rb + r) * 32 + cx + c], taps_for<8>(mx & 15, 0));
                    uint16_t *h = reinterpret_cast<uint16_t *>(&hp[i][0][0]);
                    if (r < kMctfFracPairs) h[2 * (r * kMctfFracBs + c)] = (uint16_t)u;             // pair (r, r + 1)
                    if (r >= 1) h[2 * ((r - 1) * kMctfFracBs + c) + 1] = (uint16_t)u;               // pair (r - 1, r)
                }
                __syncthreads();
                if (lane < 49) {
                    const int i = lane % 7, mx = ox + (i - 3) * step, my = oy + (lane / 7 - 3) * step;
                    if (mctf_valid(a, x, y, mx, my)) {
                        // second pass with the taps x 64 and the rounding offset x 64: clamped to
                        // [0, (1023 << 16) | 0xFFFF], the upper half of the sum is the clipped sample,
                        // and two columns' samples pack into one dword for a packed difference
                        const int8_t *vf = taps_for<8>(my & 15, 0);
                        uint32_t vt[4];
#pragma unroll
                        for (int j = 0; j < 4; j++) vt[j] = pack16(64 * vf[2 * j], 64 * vf[2 * j + 1]);
                        const int r0 = (my >> 4) - 3 - y0 - rb;      // the position's first tap row in hp
                        const uint32_t *b32 = reinterpret_cast<const uint32_t *>(L.blk);
                        int e = 0;
#pragma unroll
                        for (int r = 0; r < kMctfFracBs; r++)
#pragma unroll
                            for (int c = 0; c < kMctfFracBs; c += 2) {
                                int s0 = dot2_first(hp[i][r0 + r][c], vt[0], kVOffset << 6);
                                int s1 = dot2_first(hp[i][r0 + r][c + 1], vt[0], kVOffset << 6);
#pragma unroll
                                for (int j = 1; j < 4; j++) {
                                    s0 = dot2(hp[i][r0 + r + 2 * j][c], vt[j], s0);
                                    s1 = dot2(hp[i][r0 + r + 2 * j][c + 1], vt[j], s1);
                                }
                                const v2s d = as_v2s(b32[(r * kMctfMaxBs + c) >> 1]) -
                                              as_v2s(pack_hi16(clip_hi16(s0), clip_hi16(s1)));
                                e = dot2(as_u32(d), as_u32(d), e);
                            }
                        k = ((unsigned long long)(uint32_t)e << 32) | (uint32_t)lane;
                    }
                }
            } else if (lane < 49) {
                const int mx = ox + (lane % 7 - 3) * step, my = oy + (lane / 7 - 3) * step;
                if (mctf_valid(a, x, y, mx, my)) {
                    const int8_t *hf = taps_for<8>(mx & 15, 0), *vf = taps_for<8>(my & 15, 0);
                    const int cx = (mx >> 4) - 3 - x0, cy = (my >> 4) - 3 - y0;   // window offset of the first tap
                    uint32_t e = 0;
                    for (int c = 0; c < bs; c++) {
                        int t[kMctfMaxBs + 7];
#pragma unroll
                        for (int r = 0; r < kMctfMaxBs + 7; r++)
                            if (r < bs + 7) t[r] = hpass<8>(&L.win[(cy + r) * 32 + cx + c], hf);
                        for (int r = 0; r < bs; r++) {
                            int s = 0;
#pragma unroll
                            for (int j = 0; j < 8; j++) s += (int)vf[j] * t[r + j];
                            const int d = (int)L.blk[r * kMctfMaxBs + c] - vfinal_pix(s);
                            e += (uint32_t)(d * d);
                        }
                    }
                    k = ((unsigned long long)e << 32) | (uint32_t)lane;
                }
            }
            k = wave_min_u64(k);
            if (k != ~0ull && (uint32_t)(k >> 32) < best) {
                const int p = (int)(uint32_t)k;
                best = (uint32_t)(k >> 32);
                bmx = ox + (p % 7 - 3) * step;
                bmy = oy + (p / 7 - 3) * step;
            }
        }
    }
    if (lane == 0) {
        const intptr_t b = (intptr_t)by * a.nbx + bx;
        a.mv[b] = make_int2(bmx, bmy);
        a.err[b] = best;
    }
}

// Rounded 2x2 means of src into dst (w/2 x h/2) with dst's padding filled by
// edge replication: one thread per padded sample of dst.
__global__ void k_mctf_subsample(const uint16_t *__restrict__ src, intptr_t ss, uint16_t *__restrict__ dst,
                                 intptr_t ds, int hw, int hh, int pad)
{
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x) - pad, y = (int)blockIdx.y - pad;
    if (x >= hw + pad) return;
    const int cx = min(max(x, 0), hw - 1), cy = min(max(y, 0), hh - 1);
    const uint16_t *p = src + (intptr_t)(2 * cy) * ss + 2 * cx;
    dst[(intptr_t)y * ds + x] = (uint16_t)(((int)p[0] + p[1] + p[ss] + p[ss + 1] + 2) >> 2);
}

}  // namespace xg

using namespace xg;

namespace {

bool plane_ok(const xg_plane *p)
{
    return p && p->data && p->width >= 2 && p->height >= 2 && p->pad >= 0 && p->stride >= p->width + 2 * p->pad;
}

int32_t level_launch(hipStream_t st, const xg_plane *cur, const xg_plane *ref, int bs, int range,
                     const int32_t *lower, int lowerBw, int lowerBh, int factor, int frac, int32_t *mv, uint32_t *err)
{
    MctfArgs a;
    a.cur = cur->data;
    a.cs = cur->stride;
    a.ref = ref->data;
    a.rs = ref->stride;
    a.w = cur->width;
    a.h = cur->height;
    a.pad = ref->pad;
    a.bs = bs;
    a.range = range;
    a.frac = frac;
    a.factor = factor;
    a.lower = (const int2 *)lower;
    a.lowerBw = lowerBw;
    a.lowerBh = lowerBh;
    a.mv = (int2 *)mv;
    a.err = err;
    a.nbx = cur->width / bs;
    const int n = a.nbx * (cur->height / bs);
    if (n <= 0) return XG_OK;
    if (frac) k_mctf_level<true><<<n, 64, 0, st>>>(a);
    else k_mctf_level<false><<<n, 64, 0, st>>>(a);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

bool level_args_ok(const char *who, const xg_plane *cur, const xg_plane *ref, int bs, int range, const int32_t *lower,
                   int lowerBw, int lowerBh, int factor, const int32_t *mv, const uint32_t *err)
{
    if (!plane_ok(cur) || !plane_ok(ref) || ref->width != cur->width || ref->height != cur->height ||
        (bs != 8 && bs != 16) || range < 0 || range > kMctfMaxRange || !mv || !err ||
        (lower && (lowerBw < 1 || lowerBh < 1 || factor < 1 || factor > 4))) {
        set_error(XG_ERR_ARG, "%s: bad arguments (planes, blockSize 8 or 16, range 0..%d, lower grid, factor 1..4, "
                              "outputs)", who, kMctfMaxRange);
        return false;
    }
    return true;
}

// Level pictures and fields of xg_mctf_me, one set per session.
struct MctfBufs {
    uint16_t *pic = nullptr;
    int32_t *mv = nullptr;
    uint32_t *err = nullptr;
    size_t picCap = 0, mvCap = 0, errCap = 0;
};
std::mutex g_mctf_mu;
std::vector<std::pair<xg_session *, MctfBufs>> g_mctf;

template <class T>
bool grow(T *&p, size_t &cap, size_t n, hipStream_t st)
{
    if (n <= cap) return true;
    if (p) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    if (!XG_HIP(hipMalloc((void **)&p, n * sizeof(T)))) return false;
    cap = n;
    return true;
}

}  // namespace

namespace xg {
void mctf_release(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_mctf_mu);
    for (size_t i = 0; i < g_mctf.size(); i++)
        if (g_mctf[i].first == s) {
            (void)hipFree(g_mctf[i].second.pic);
            (void)hipFree(g_mctf[i].second.mv);
            (void)hipFree(g_mctf[i].second.err);
            g_mctf.erase(g_mctf.begin() + (long)i);
            return;
        }
}
}  // namespace xg

extern "C" {

int32_t xg_mctf_subsample(xg_session *s, const xg_plane *src, const xg_plane *dst)
{
    SessionView v;
    if (!s || !session_view(s, &v) || !plane_ok(src) || !plane_ok(dst) || dst->width != src->width / 2 ||
        dst->height != src->height / 2) {
        set_error(XG_ERR_ARG, "xg_mctf_subsample: bad arguments (planes; dst must be src / 2 in both sides)");
        return XG_ERR_ARG;
    }
    const int pad = dst->pad, hw = dst->width, hh = dst->height;
    k_mctf_subsample<<<dim3((hw + 2 * pad + 255) / 256, hh + 2 * pad), 256, 0, v.stream>>>(
        src->data, src->stride, dst->data, dst->stride, hw, hh, pad);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

int32_t xg_mctf_me_level(xg_session *s, const xg_plane *cur, const xg_plane *ref, int32_t blockSize, int32_t range,
                         const int32_t *lower, int32_t lowerBw, int32_t lowerBh, int32_t factor, int32_t frac,
                         int32_t *mv, uint32_t *err)
{
    SessionView v;
    if (!s || !session_view(s, &v)) {
        set_error(XG_ERR_ARG, "xg_mctf_me_level: bad arguments (no session)");
        return XG_ERR_ARG;
    }
    if (!level_args_ok("xg_mctf_me_level", cur, ref, blockSize, range, lower, lowerBw, lowerBh, factor, mv, err))
        return XG_ERR_ARG;
    return level_launch(v.stream, cur, ref, blockSize, range, lower, lowerBw, lowerBh, factor, frac ? 1 : 0, mv, err);
}

int32_t xg_mctf_me(xg_session *s, const xg_plane *cur, const xg_plane *ref, int32_t *mv, uint32_t *err)
{
    SessionView v;
    if (!s || !session_view(s, &v) || !plane_ok(cur) || !plane_ok(ref) || ref->width != cur->width ||
        ref->height != cur->height || cur->width < 64 || cur->height < 64 || !mv || !err) {
        set_error(XG_ERR_ARG, "xg_mctf_me: bad arguments (planes of the same size, at least 64x64; outputs)");
        return XG_ERR_ARG;
    }
    const int W = cur->width, H = cur->height;
    const int w2 = W / 2, h2 = H / 2, w4 = w2 / 2, h4 = h2 / 2;
    const int p2 = 32, p4 = 16;
    const size_t n2 = (size_t)(w2 + 2 * p2) * (h2 + 2 * p2), n4 = (size_t)(w4 + 2 * p4) * (h4 + 2 * p4);
    const size_t c2 = (size_t)w2 * h2, c4 = (size_t)w4 * h4;
    const int g4x = w4 / 16, g4y = h4 / 16, g2x = w2 / 16, g2y = h2 / 16, g1x = W / 16, g1y = H / 16;
    const size_t nmv = (size_t)g4x * g4y + (size_t)g2x * g2y + (size_t)g1x * g1y;
    MctfBufs *b = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mctf_mu);
        for (auto &p : g_mctf)
            if (p.first == s) b = &p.second;
        if (!b) {
            g_mctf.push_back({s, MctfBufs()});
            b = &g_mctf.back().second;
        }
        if (!grow(b->pic, b->picCap, n2 + n4 + c2 + c4, v.stream) || !grow(b->mv, b->mvCap, 2 * nmv, v.stream) ||
            !grow(b->err, b->errCap, nmv, v.stream))
            return XG_ERR_NOMEM;
    }
    uint16_t *r2 = b->pic, *r4 = r2 + n2, *q2 = r4 + n4, *q4 = q2 + c2;
    const xg_plane R2{r2 + (size_t)p2 * (w2 + 2 * p2) + p2, w2 + 2 * p2, w2, h2, p2};
    const xg_plane R4{r4 + (size_t)p4 * (w4 + 2 * p4) + p4, w4 + 2 * p4, w4, h4, p4};
    const xg_plane C2{q2, w2, w2, h2, 0}, C4{q4, w4, w4, h4, 0};
    int32_t *m4 = b->mv, *m2 = m4 + 2 * (size_t)g4x * g4y, *m1 = m2 + 2 * (size_t)g2x * g2y;
    uint32_t *e4 = b->err, *e2 = e4 + (size_t)g4x * g4y, *e1 = e2 + (size_t)g2x * g2y;
    int32_t rc = XG_OK;
    int evFrame[2], evLast[2];
    const bool tFrame = session_timer_begin(s, 9, evFrame);
    if (rc == XG_OK) rc = xg_mctf_subsample(s, cur, &C2);
    if (rc == XG_OK) rc = xg_mctf_subsample(s, &C2, &C4);
    if (rc == XG_OK) rc = xg_mctf_subsample(s, ref, &R2);
    if (rc == XG_OK) rc = xg_mctf_subsample(s, &R2, &R4);
    if (rc == XG_OK) rc = level_launch(v.stream, &C4, &R4, 16, 8, nullptr, 0, 0, 1, 0, m4, e4);
    if (rc == XG_OK) rc = level_launch(v.stream, &C2, &R2, 16, 5, m4, g4x, g4y, 2, 0, m2, e2);
    if (rc == XG_OK) rc = level_launch(v.stream, cur, ref, 16, 5, m2, g2x, g2y, 2, 0, m1, e1);
    const bool tLast = session_timer_begin(s, 10, evLast);
    if (rc == XG_OK) rc = level_launch(v.stream, cur, ref, 8, 0, m1, g1x, g1y, 1, 1, mv, err);
    if (tLast) session_timer_end(s, 10, evLast);
    if (tFrame) session_timer_end(s, 9, evFrame);
    return rc;
}

}  // extern "C"

// xg_refine.hip -- batched fractional motion refinement + motion compensation.
//
// Stage 1 (k_refine): for every block and each of its ncand candidate MVs
// (base MV + offset, 1/16 pel, clipped to the CU MV window of
// h266_encoder/source/h266_enc_init.c:1462-1465), the luma 8-tap prediction
// (h266_core/source/h266_luma_mc.c, table dispatch of
// h266_encoder/source/h266_inter_prediction.c:143-172, filterIndex 0) and,
// when enabled, the Cb/Cr 4-tap predictions (h266_chroma_mc.c, address split
// of h266_inter_prediction.c:664-674) are formed and compared with the
// current block by SATD; cost = (SATD << 8) + (EG(dx) + EG(dy)) * lambda with
// the MVD in 1/16 pel.
// Stage 2 (k_mc): per block the lowest cost (earliest candidate on ties), its
// result record, and the luma + chroma prediction of that candidate written
// into the prediction picture.
//
// Mapping (one wavefront = one workgroup of 64 lanes):
//  * a task is (block, first candidate, candidate count); tasks are grouped
//    per block shape and ordered largest first (per-shape work queues);
//  * the task's reference window -- every sample any of its candidates'
//    filters touch -- is staged once into LDS and reused by all candidates;
//  * lane = candidate * W + column: each lane walks its column down the block,
//    keeping the 8 (luma) / 4 (chroma) first-pass values of the separable
//    filter in registers, so the intermediate never leaves the VGPRs;
//  * SATD tiles (8x8, 4x4 or 2x2 as the block shape allows) are transformed
//    vertically in registers and horizontally across lanes with DPP and
//    ds_swizzle butterflies; sums reduce within the wavefront;
//  * batched mode: one persistent launch per stage pulls tasks from an atomic
//    counter; plain mode: one launch per block shape.
#include <algorithm>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

constexpr int kLumaLW = 82;        // LDS row stride (u16) of the luma window: 41 words
constexpr int kLumaLH = 80;        // rows: H + 7 + 9 at most
constexpr int kChromaLW = 42;      // per-plane chroma window stride (u16): 21 words
constexpr int kChromaLH = 40;      // rows: H/2 + 3 + 5 at most
constexpr int kMaxOff = 64;        // |candidate offset| limit, 1/16 pel
constexpr int kNumShapes = 25;     // (lgW-2)*5 + (lgH-2), W, H in 4..64
constexpr int kPicSlots = 16;

struct RefineArgs {
    const xg_picture *pics;
    const xg_block *blocks;
    const int2 *tasks;             // x: block index, y: c0 | (count << 16)
    int taskBegin, taskEnd;        // plain mode: this launch's range
    int *counter;                  // batched mode: work-queue head (nullptr in plain mode)
    uint32_t *candCost;
    int ncand;
    int offMinX, offMaxX, offMinY, offMaxY;
    uint32_t lambda;
    int ctuSize;
    int useChroma;
    const int16_t *off;            // device: offx[XG_MAX_CAND] then offy[XG_MAX_CAND]
};

struct McArgs {
    const xg_picture *pics;
    const xg_block *blocks;
    const int *order;              // block indices grouped by shape
    int begin, end;
    int *counter;
    const uint32_t *candCost;
    xg_block_result *best;
    xg_picture pred;               // output planes (unpadded)
    int ncand;
    int ctuSize;
    const int16_t *off;
};

// CU MV window in 1/16 pel (h266_enc_init.c:1462-1465, int16 as stored there).
struct MvWin { int mnx, mny, mxx, mxy; };
__device__ __forceinline__ MvWin mv_window(const xg_block &B, int fw, int fh, int ctu)
{
    MvWin w;
    w.mnx = (int)(int16_t)(-B.x - ctu - 8) * 16;
    w.mny = (int)(int16_t)(-B.y - ctu - 8) * 16;
    w.mxx = (int)(int16_t)(fw + (ctu - B.w + 8) - B.x) * 16;
    w.mxy = (int)(int16_t)(fh + (ctu - B.h + 8) - B.y) * 16;
    return w;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

template <int TAPS>
__device__ __forceinline__ int hrow_lds(const uint16_t *p, const int *hc)
{
    int s = 0;
#pragma unroll
    for (int j = 0; j < TAPS; j++) s += (int)p[j] * hc[j];
    return (int)(int16_t)((s + kHOffset) >> kHShift);
}

// Luma SATD of one pass of candidates; lane = cand_local * W + x.
template <int W, int TS>
__device__ __forceinline__ void luma_pass(const RefineArgs &a, const xg_block &B, const MvWin &mw,
                                          const uint16_t *lds, int wx0, int wy0,
                                          const uint16_t *cur, intptr_t cs, int c, bool valid, int ci,
                                          uint32_t *csum, const int16_t *sOff)
{
    const int lane = threadIdx.x;
    const int xl = lane & (W - 1);
    const int H = B.h;
    const int mvx = clampi(B.mvx + sOff[c], mw.mnx, mw.mxx);
    const int mvy = clampi(B.mvy + sOff[XG_MAX_CAND + c], mw.mny, mw.mxy);
    const int rx = (B.x << 4) + mvx, ry = (B.y << 4) + mvy;
    int hc[8], vc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        hc[j] = c_luma[0][rx & 15][j];
        vc[j] = c_luma[0][ry & 15][j];
    }
    const uint16_t *wp = lds + ((ry >> 4) - 3 - wy0) * kLumaLW + ((rx >> 4) - 3 - wx0) + xl;
    const uint16_t *cp = cur + xl;
    int t[TS + 7];
#pragma unroll
    for (int k = 0; k < 7; k++) t[k] = hrow_lds<8>(wp + k * kLumaLW, hc);
    int acc = 0;
    for (int y0 = 0; y0 < H; y0 += TS) {
#pragma unroll
        for (int k = 0; k < TS; k++) t[7 + k] = hrow_lds<8>(wp + (y0 + 7 + k) * kLumaLW, hc);
        int d[TS];
#pragma unroll
        for (int k = 0; k < TS; k++) {
            int s = 0;
#pragma unroll
            for (int j = 0; j < 8; j++) s += vc[j] * t[k + j];
            d[k] = (int)cp[(intptr_t)(y0 + k) * cs] - vfinal_pix(s);
        }
        const int ts = tile_satd<TS>(d, lane);
        if ((xl & (TS - 1)) == 0) acc += ts;
#pragma unroll
        for (int k = 0; k < 7; k++) t[k] = t[TS + k];
    }
    acc = group_sum<W>(acc);
    if (xl == 0 && valid) csum[ci] = (uint32_t)acc;
}

// Chroma SATD (Cb + Cr) of one pass; lane = cand_local * WC + x.
template <int WC, int TC>
__device__ __forceinline__ void chroma_pass(const RefineArgs &a, const xg_block &B, const MvWin &mw,
                                            const uint16_t *ldsU, const uint16_t *ldsV, int cx0, int cy0,
                                            const uint16_t *curU, const uint16_t *curV, intptr_t cs,
                                            int c, bool valid, int ci, uint32_t *csum, const int16_t *sOff)
{
    const int lane = threadIdx.x;
    const int xl = lane & (WC - 1);
    const int HC = B.h >> 1;
    const int mvx = clampi(B.mvx + sOff[c], mw.mnx, mw.mxx);
    const int mvy = clampi(B.mvy + sOff[XG_MAX_CAND + c], mw.mny, mw.mxy);
    const int rx = (B.x << 4) + mvx, ry = (B.y << 4) + mvy;
    int hc[4], vc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        hc[j] = c_chroma[rx & 31][j];
        vc[j] = c_chroma[ry & 31][j];
    }
    const int off = ((ry >> 5) - 1 - cy0) * kChromaLW + ((rx >> 5) - 1 - cx0) + xl;
    int acc = 0;
#pragma unroll
    for (int pl = 0; pl < 2; pl++) {
        const uint16_t *wp = (pl ? ldsV : ldsU) + off;
        const uint16_t *cp = (pl ? curV : curU) + xl;
        int t[TC + 3];
#pragma unroll
        for (int k = 0; k < 3; k++) t[k] = hrow_lds<4>(wp + k * kChromaLW, hc);
        for (int y0 = 0; y0 < HC; y0 += TC) {
#pragma unroll
            for (int k = 0; k < TC; k++) t[3 + k] = hrow_lds<4>(wp + (y0 + 3 + k) * kChromaLW, hc);
            int d[TC];
#pragma unroll
            for (int k = 0; k < TC; k++) {
                int s = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) s += vc[j] * t[k + j];
                d[k] = (int)cp[(intptr_t)(y0 + k) * cs] - vfinal_pix(s);
            }
            const int ts = tile_satd<TC>(d, lane);
            if ((xl & (TC - 1)) == 0) acc += ts;
#pragma unroll
            for (int k = 0; k < 3; k++) t[k] = t[TC + k];
        }
    }
    acc = group_sum<WC>(acc);
    if (xl == 0 && valid) csum[ci] += (uint32_t)acc;
}

template <int W, int TS>
__device__ __forceinline__ void luma_task(const RefineArgs &a, const xg_block &B, const MvWin &mw, const uint16_t *lds,
                          int wx0, int wy0, const uint16_t *cur, intptr_t cs, int c0, int nc, uint32_t *csum,
                          const int16_t *sOff)
{
    constexpr int CP = 64 / W;
    const int cl = threadIdx.x / W;
    for (int base = 0; base < nc; base += CP) {
        const int ci = base + cl;
        const bool valid = ci < nc;
        luma_pass<W, TS>(a, B, mw, lds, wx0, wy0, cur, cs, c0 + min(ci, nc - 1), valid, ci, csum, sOff);
    }
}

template <int WC, int TC>
__device__ __forceinline__ void chroma_task(const RefineArgs &a, const xg_block &B, const MvWin &mw, const uint16_t *ldsU,
                            const uint16_t *ldsV, int cx0, int cy0, const uint16_t *curU,
                            const uint16_t *curV, intptr_t cs, int c0, int nc, uint32_t *csum,
                            const int16_t *sOff)
{
    constexpr int CP = 64 / WC;
    const int cl = threadIdx.x / WC;
    for (int base = 0; base < nc; base += CP) {
        const int ci = base + cl;
        const bool valid = ci < nc;
        chroma_pass<WC, TC>(a, B, mw, ldsU, ldsV, cx0, cy0, curU, curV, cs, c0 + min(ci, nc - 1), valid,
                            ci, csum, sOff);
    }
}

__device__ __forceinline__ void refine_one(const RefineArgs &a, int task, uint16_t *lds, uint32_t *csum, const int16_t *sOff)
{
    const int lane = threadIdx.x;
    const int2 tk = a.tasks[task];
    const int b = tk.x, c0 = tk.y & 0xFFFF, nc = tk.y >> 16;
    const xg_block B = a.blocks[b];
    const xg_picture &P = a.pics[B.pic];
    const xg_picture &R = a.pics[B.ref];
    const MvWin mw = mv_window(B, P.plane[0].width, P.plane[0].height, a.ctuSize);

    // ---- luma window: every sample the task's candidates read ----
    const int ixmin = ((B.x << 4) + clampi(B.mvx + a.offMinX, mw.mnx, mw.mxx)) >> 4;
    const int ixmax = ((B.x << 4) + clampi(B.mvx + a.offMaxX, mw.mnx, mw.mxx)) >> 4;
    const int iymin = ((B.y << 4) + clampi(B.mvy + a.offMinY, mw.mny, mw.mxy)) >> 4;
    const int iymax = ((B.y << 4) + clampi(B.mvy + a.offMaxY, mw.mny, mw.mxy)) >> 4;
    const int wx0 = ixmin - 3, wy0 = iymin - 3;
    const int WW = ixmax - ixmin + B.w + 7, HH = iymax - iymin + B.h + 7;
    {
        const intptr_t rs = R.plane[0].stride;
        const uint16_t *rp = R.plane[0].data + (intptr_t)wy0 * rs + wx0;
        for (int r = 0; r < HH; r++)
            for (int cc = lane; cc < WW; cc += 64) lds[r * kLumaLW + cc] = rp[(intptr_t)r * rs + cc];
    }
    csum[lane] = 0;
    __syncthreads();

    const intptr_t cs = P.plane[0].stride;
    const uint16_t *cur = P.plane[0].data + (intptr_t)B.y * cs + B.x;
    const bool t4 = B.w == 4 || B.h == 4;
    switch (B.w) {
    case 4: luma_task<4, 4>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff); break;
    case 8: t4 ? luma_task<8, 4>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff)
               : luma_task<8, 8>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff); break;
    case 16: t4 ? luma_task<16, 4>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff)
                : luma_task<16, 8>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff); break;
    case 32: t4 ? luma_task<32, 4>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff)
                : luma_task<32, 8>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff); break;
    default: t4 ? luma_task<64, 4>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff)
                : luma_task<64, 8>(a, B, mw, lds, wx0, wy0, cur, cs, c0, nc, csum, sOff); break;
    }
    __syncthreads();

    if (a.useChroma) {
        // ---- chroma windows (Cb, Cr share the address split) ----
        const int cxmin = ((B.x << 4) + clampi(B.mvx + a.offMinX, mw.mnx, mw.mxx)) >> 5;
        const int cxmax = ((B.x << 4) + clampi(B.mvx + a.offMaxX, mw.mnx, mw.mxx)) >> 5;
        const int cymin = ((B.y << 4) + clampi(B.mvy + a.offMinY, mw.mny, mw.mxy)) >> 5;
        const int cymax = ((B.y << 4) + clampi(B.mvy + a.offMaxY, mw.mny, mw.mxy)) >> 5;
        const int cx0 = cxmin - 1, cy0 = cymin - 1;
        const int CWW = cxmax - cxmin + (B.w >> 1) + 3, CHH = cymax - cymin + (B.h >> 1) + 3;
        uint16_t *ldsU = lds, *ldsV = lds + kChromaLW * kChromaLH;
        const intptr_t rs = R.plane[1].stride;
        const uint16_t *ru = R.plane[1].data + (intptr_t)cy0 * rs + cx0;
        const uint16_t *rv = R.plane[2].data + (intptr_t)cy0 * rs + cx0;
        for (int r = 0; r < CHH; r++)
            for (int cc = lane; cc < 2 * CWW; cc += 64) {
                const int pl = cc >= CWW, col = cc - pl * CWW;
                (pl ? ldsV : ldsU)[r * kChromaLW + col] = (pl ? rv : ru)[(intptr_t)r * rs + col];
            }
        __syncthreads();
        const intptr_t ccs = P.plane[1].stride;
        const intptr_t co = (intptr_t)(B.y >> 1) * ccs + (B.x >> 1);
        const uint16_t *cu = P.plane[1].data + co, *cv = P.plane[2].data + co;
        const int wc = B.w >> 1, hc = B.h >> 1;
        const int tc = (wc % 8 == 0 && hc % 8 == 0) ? 8 : (wc % 4 == 0 && hc % 4 == 0) ? 4 : 2;
        switch (wc) {
        case 2: chroma_task<2, 2>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff); break;
        case 4: tc == 4 ? chroma_task<4, 4>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff)
                        : chroma_task<4, 2>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff); break;
        case 8: tc == 8 ? chroma_task<8, 8>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff)
              : tc == 4 ? chroma_task<8, 4>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff)
                        : chroma_task<8, 2>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff); break;
        case 16: tc == 8 ? chroma_task<16, 8>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff)
               : tc == 4 ? chroma_task<16, 4>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff)
                         : chroma_task<16, 2>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff); break;
        default: tc == 8 ? chroma_task<32, 8>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff)
               : tc == 4 ? chroma_task<32, 4>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff)
                         : chroma_task<32, 2>(a, B, mw, ldsU, ldsV, cx0, cy0, cu, cv, ccs, c0, nc, csum, sOff); break;
        }
        __syncthreads();
    }

    if (lane < nc) {
        const int c = c0 + lane;
        const int mvx = clampi(B.mvx + sOff[c], mw.mnx, mw.mxx);
        const int mvy = clampi(B.mvy + sOff[XG_MAX_CAND + c], mw.mny, mw.mxy);
        const uint32_t bits = eg_bits(mvx - B.mvpx) + eg_bits(mvy - B.mvpy);
        a.candCost[(size_t)b * a.ncand + c] = (csum[lane] << 8) + bits * a.lambda;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) k_refine(RefineArgs a)
{
    __shared__ uint16_t lds[kLumaLW * kLumaLH];
    __shared__ uint32_t csum[64];
    __shared__ int16_t sOff[2 * XG_MAX_CAND];
    __shared__ int s_task;
    for (int i = threadIdx.x; i < 2 * XG_MAX_CAND; i += 64) sOff[i] = a.off[i];
    __syncthreads();
    if (a.counter) {
        for (;;) {
            if (threadIdx.x == 0) s_task = atomicAdd(a.counter, 1);
            __syncthreads();
            const int t = s_task;
            __syncthreads();
            if (t >= a.taskEnd) break;
            refine_one(a, t, lds, csum, sOff);
        }
    } else {
        const int t = a.taskBegin + (int)blockIdx.x;
        if (t < a.taskEnd) refine_one(a, t, lds, csum, sOff);
    }
}

// ---------------------------------------------------------------------------
// Stage 2: argmin + motion compensation of the best candidate.
// ---------------------------------------------------------------------------
template <int TAPS, int FBITS>
__device__ __forceinline__ void mc_column(const uint16_t *src, intptr_t ss, uint16_t *dst, intptr_t ds,
                                          int fx, int fy, int h)
{
    // src: top-left of the integer position of this column's block; the
    // column walks down h rows with a sliding window of first-pass values.
    const int8_t *ch = taps_for<TAPS>(fx, 0);
    const int8_t *cv = taps_for<TAPS>(fy, 0);
    int hc[TAPS], vc[TAPS];
#pragma unroll
    for (int j = 0; j < TAPS; j++) {
        hc[j] = ch[j];
        vc[j] = cv[j];
    }
    const uint16_t *p = src - (TAPS / 2 - 1) * ss - (TAPS / 2 - 1);
    int t[TAPS];
#pragma unroll
    for (int k = 0; k < TAPS - 1; k++) t[k] = hrow_lds<TAPS>(p + k * ss, hc);
    for (int y = 0; y < h; y++) {
        t[TAPS - 1] = hrow_lds<TAPS>(p + (y + TAPS - 1) * ss, hc);
        int s = 0;
#pragma unroll
        for (int j = 0; j < TAPS; j++) s += vc[j] * t[j];
        dst[(intptr_t)y * ds] = (uint16_t)vfinal_pix(s);
#pragma unroll
        for (int k = 0; k < TAPS - 1; k++) t[k] = t[k + 1];
    }
}

__device__ __forceinline__ void mc_one(const McArgs &a, int b, const int16_t *sOff)
{
    const int lane = threadIdx.x;
    const xg_block B = a.blocks[b];
    const xg_picture &P = a.pics[B.pic];
    const xg_picture &R = a.pics[B.ref];
    const MvWin mw = mv_window(B, P.plane[0].width, P.plane[0].height, a.ctuSize);
    uint32_t cost = lane < a.ncand ? a.candCost[(size_t)b * a.ncand + lane] : 0xFFFFFFFFu;
    int idx = lane;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const uint32_t oc = (uint32_t)__shfl_xor((int)cost, m, 64);
        const int oi = __shfl_xor(idx, m, 64);
        if (oc < cost || (oc == cost && oi < idx)) {
            cost = oc;
            idx = oi;
        }
    }
    const int mvx = clampi(B.mvx + sOff[idx], mw.mnx, mw.mxx);
    const int mvy = clampi(B.mvy + sOff[XG_MAX_CAND + idx], mw.mny, mw.mxy);
    if (lane == 0) {
        xg_block_result r;
        r.mvx = mvx;
        r.mvy = mvy;
        r.cost = cost;
        r.cand = idx;
        a.best[b] = r;
    }
    const int rx = (B.x << 4) + mvx, ry = (B.y << 4) + mvy;
    if (lane < B.w) {
        const intptr_t rs = R.plane[0].stride;
        const uint16_t *src = R.plane[0].data + (intptr_t)(ry >> 4) * rs + (rx >> 4) + lane;
        uint16_t *dst = a.pred.plane[0].data + (intptr_t)B.y * a.pred.plane[0].stride + B.x + lane;
        mc_column<8, 4>(src, rs, dst, a.pred.plane[0].stride, rx & 15, ry & 15, B.h);
    }
    const int wc = B.w >> 1;
    if (lane < 2 * wc) {
        const int pl = 1 + (lane >= wc), xl = lane - (pl - 1) * wc;
        const intptr_t rs = R.plane[pl].stride;
        const uint16_t *src = R.plane[pl].data + (intptr_t)(ry >> 5) * rs + (rx >> 5) + xl;
        uint16_t *dst = a.pred.plane[pl].data + (intptr_t)(B.y >> 1) * a.pred.plane[pl].stride + (B.x >> 1) + xl;
        mc_column<4, 5>(src, rs, dst, a.pred.plane[pl].stride, rx & 31, ry & 31, B.h >> 1);
    }
}

__global__ void __launch_bounds__(64) k_mc(McArgs a)
{
    __shared__ int s_idx;
    __shared__ int16_t sOff[2 * XG_MAX_CAND];
    for (int i = threadIdx.x; i < 2 * XG_MAX_CAND; i += 64) sOff[i] = a.off[i];
    __syncthreads();
    if (a.counter) {
        for (;;) {
            if (threadIdx.x == 0) s_idx = atomicAdd(a.counter, 1);
            __syncthreads();
            const int i = s_idx;
            __syncthreads();
            if (i >= a.end) break;
            mc_one(a, a.order[i], sOff);
        }
    } else {
        const int i = a.begin + (int)blockIdx.x;
        if (i < a.end) mc_one(a, a.order[i], sOff);
    }
}

__global__ void k_zero_counters(int *c, int n)
{
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

}  // namespace xg

// ---------------------------------------------------------------------------
// Session
// ---------------------------------------------------------------------------
using namespace xg;

struct xg_session {
    hipStream_t stream = nullptr;
    int maxBlocks = 0;
    int numCUs = 256;
    int persistFactor = 16;            // persistent workgroups per CU

    xg_picture pics[kPicSlots];
    bool picSet[kPicSlots] = {};
    xg_picture *dPics = nullptr;

    std::vector<xg_block> blocks;
    xg_block *dBlocks = nullptr;
    int nblocks = 0;
    uint64_t blocksVersion = 1;
    int lastNcand = 0;                 // ncand of the last refinement

    // per-shape work queues
    int tile[kNumShapes];              // candidates per task
    std::vector<int2> tasks;
    int2 *dTasks = nullptr;
    size_t taskCap = 0;
    int shapeTaskBegin[kNumShapes + 1] = {};
    std::vector<int> shapeOrder;       // shapes, heaviest tasks first
    int *dOrder = nullptr;             // block indices grouped by shape
    int shapeBlockBegin[kNumShapes + 1] = {};
    uint64_t tasksKey = 0;             // (blocksVersion, ncand, tile table) the tasks were built for

    uint32_t *dCandCost = nullptr;
    xg_block_result *dBest = nullptr;
    int16_t *dOff = nullptr;           // candidate offsets of the last configuration
    int16_t hOff[2 * XG_MAX_CAND];
    bool offValid = false;
    int *dCounters = nullptr;
    void *hStage = nullptr;            // pinned upload staging
    size_t hStageCap = 0;

    // graph of the batched stages
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    uint64_t graphKey = 0;

    // timing
    bool timing = false;
    std::vector<hipEvent_t> evPool;
    std::vector<std::pair<int, int>> evPending[2];  // per stage: (start, stop) indices
    int evNext = 0;
    double accMs[2] = {0, 0};
    int accN[2] = {0, 0};
};

namespace xg {

static int shape_of(int w, int h)
{
    int lw = 0, lh = 0;
    while ((4 << lw) < w) lw++;
    while ((4 << lh) < h) lh++;
    return lw * 5 + lh;
}

static bool stage_upload(xg_session *s, void *dst, const void *src, size_t bytes)
{
    if (bytes > s->hStageCap) {
        if (s->hStage) {
            if (!XG_HIP(hipStreamSynchronize(s->stream))) return false;
            hipHostFree(s->hStage);
        }
        s->hStage = nullptr;
        s->hStageCap = 0;
        size_t cap = align_up(bytes, 1 << 16);
        if (!XG_HIP(hipHostMalloc(&s->hStage, cap, hipHostMallocDefault))) return false;
        s->hStageCap = cap;
    }
    // the staging buffer is reused: wait for earlier uploads from it
    if (!XG_HIP(hipStreamSynchronize(s->stream))) return false;
    memcpy(s->hStage, src, bytes);
    return XG_HIP(hipMemcpyAsync(dst, s->hStage, bytes, hipMemcpyHostToDevice, s->stream));
}

static uint64_t mix(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h;
}

static bool build_tasks(xg_session *s, int ncand)
{
    uint64_t key = mix(mix(s->blocksVersion, (uint64_t)ncand), 0);
    for (int i = 0; i < kNumShapes; i++) key = mix(key, (uint64_t)s->tile[i]);
    if (key == s->tasksKey) return true;
    std::vector<std::vector<int>> per(kNumShapes);
    for (int b = 0; b < s->nblocks; b++) per[shape_of(s->blocks[b].w, s->blocks[b].h)].push_back(b);
    // order shapes by task weight (samples per task), heaviest first
    std::vector<int> order(kNumShapes);
    for (int i = 0; i < kNumShapes; i++) order[i] = i;
    auto weight = [&](int sh) {
        const int w = 4 << (sh / 5), h = 4 << (sh % 5);
        return (long)w * h * std::min(s->tile[sh], ncand);
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return weight(x) > weight(y); });
    s->tasks.clear();
    std::vector<int> blockOrder;
    blockOrder.reserve(s->nblocks);
    int sb[kNumShapes + 1], st[kNumShapes + 1];
    for (int oi = 0; oi < kNumShapes; oi++) {
        const int sh = order[oi];
        st[oi] = (int)s->tasks.size();
        sb[oi] = (int)blockOrder.size();
        const int cpt = std::max(1, std::min(s->tile[sh], ncand));
        for (int b : per[sh]) {
            blockOrder.push_back(b);
            for (int c0 = 0; c0 < ncand; c0 += cpt) s->tasks.push_back(make_int2(b, c0 | (std::min(cpt, ncand - c0) << 16)));
        }
    }
    st[kNumShapes] = (int)s->tasks.size();
    sb[kNumShapes] = (int)blockOrder.size();
    memcpy(s->shapeTaskBegin, st, sizeof(st));
    memcpy(s->shapeBlockBegin, sb, sizeof(sb));
    s->shapeOrder = order;
    if (s->tasks.size() > s->taskCap) {
        if (s->dTasks) {
            if (!XG_HIP(hipStreamSynchronize(s->stream))) return false;
            hipFree(s->dTasks);
        }
        s->dTasks = nullptr;
        s->taskCap = 0;
        size_t cap = std::max<size_t>(s->tasks.size(), 1024);
        if (!XG_HIP(hipMalloc((void **)&s->dTasks, cap * sizeof(int2)))) return false;
        s->taskCap = cap;
    }
    if (!s->tasks.empty() && !stage_upload(s, s->dTasks, s->tasks.data(), s->tasks.size() * sizeof(int2))) return false;
    if (!blockOrder.empty() && !stage_upload(s, s->dOrder, blockOrder.data(), blockOrder.size() * sizeof(int))) return false;
    if (!XG_HIP(hipStreamSynchronize(s->stream))) return false;
    s->tasksKey = key;
    return true;
}

static int ev_get(xg_session *s)
{
    if (s->evNext >= (int)s->evPool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        s->evPool.push_back(e);
    }
    return s->evNext++;
}

static void ev_collect(xg_session *s)
{
    for (int st = 0; st < 2; st++) {
        for (auto &pr : s->evPending[st]) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, s->evPool[pr.first], s->evPool[pr.second]) == hipSuccess) {
                s->accMs[st] += ms;
                s->accN[st]++;
            }
        }
        s->evPending[st].clear();
    }
    s->evNext = 0;
}

static bool validate_cfg(const xg_refine_cfg *cfg)
{
    if (!cfg || cfg->ncand < 1 || cfg->ncand > XG_MAX_CAND || cfg->ctuSize < 4 || cfg->ctuSize > 256) {
        set_error(XG_ERR_ARG, "xg_refine_frac: bad configuration");
        return false;
    }
    for (int i = 0; i < cfg->ncand; i++)
        if (abs(cfg->offx[i]) > kMaxOff || abs(cfg->offy[i]) > kMaxOff) {
            set_error(XG_ERR_ARG, "xg_refine_frac: candidate offset beyond +/-%d (1/16 pel)", kMaxOff);
            return false;
        }
    return true;
}

static bool validate_blocks(xg_session *s, const xg_picture *pred, int ctu)
{
    for (int i = 0; i < s->nblocks; i++) {
        const xg_block &B = s->blocks[i];
        if (B.pic >= kPicSlots || B.ref >= kPicSlots || !s->picSet[B.pic] || !s->picSet[B.ref]) {
            set_error(XG_ERR_ARG, "block %d: picture slot not set", i);
            return false;
        }
        const xg_picture &P = s->pics[B.pic], &R = s->pics[B.ref];
        if (B.x < 0 || B.y < 0 || B.x + B.w > P.plane[0].width || B.y + B.h > P.plane[0].height ||
            (B.x & 1) || (B.y & 1)) {
            set_error(XG_ERR_ARG, "block %d (%d,%d %dx%d) outside its picture", i, B.x, B.y, B.w, B.h);
            return false;
        }
        if (R.plane[0].width != P.plane[0].width || R.plane[0].height != P.plane[0].height ||
            R.plane[0].pad < ctu + 16 || R.plane[1].pad < ctu / 2 + 8 || R.plane[2].pad < ctu / 2 + 8) {
            set_error(XG_ERR_ARG, "block %d: reference picture size/padding (needs luma pad >= ctuSize+16)", i);
            return false;
        }
    }
    if (pred)
        for (int i = 0; i < s->nblocks; i++) {
            const xg_block &B = s->blocks[i];
            for (int p = 0; p < 3; p++) {
                const int sub = p ? 1 : 0;
                const xg_plane &pl = pred->plane[p];
                if (!pl.data || pl.stride < pl.width || ((B.x + B.w) >> sub) > pl.width ||
                    ((B.y + B.h) >> sub) > pl.height) {
                    set_error(XG_ERR_ARG, "prediction plane %d does not cover block %d", p, i);
                    return false;
                }
            }
        }
    return true;
}

static bool fill_args(xg_session *s, const xg_refine_cfg *cfg, RefineArgs &ra, McArgs &ma, const xg_picture *pred)
{
    int16_t off[2 * XG_MAX_CAND] = {};
    memset(&ra, 0, sizeof(ra));
    ra.pics = s->dPics;
    ra.blocks = s->dBlocks;
    ra.tasks = s->dTasks;
    ra.candCost = s->dCandCost;
    ra.ncand = cfg->ncand;
    ra.offMinX = ra.offMinY = 1 << 30;
    ra.offMaxX = ra.offMaxY = -(1 << 30);
    for (int i = 0; i < cfg->ncand; i++) {
        off[i] = cfg->offx[i];
        off[XG_MAX_CAND + i] = cfg->offy[i];
        ra.offMinX = std::min(ra.offMinX, (int)cfg->offx[i]);
        ra.offMaxX = std::max(ra.offMaxX, (int)cfg->offx[i]);
        ra.offMinY = std::min(ra.offMinY, (int)cfg->offy[i]);
        ra.offMaxY = std::max(ra.offMaxY, (int)cfg->offy[i]);
    }
    ra.lambda = cfg->lambda;
    ra.ctuSize = cfg->ctuSize;
    ra.useChroma = cfg->useChroma;
    ra.off = s->dOff;
    memset(&ma, 0, sizeof(ma));
    ma.pics = s->dPics;
    ma.blocks = s->dBlocks;
    ma.order = s->dOrder;
    ma.candCost = s->dCandCost;
    ma.best = s->dBest;
    if (pred) ma.pred = *pred;
    ma.ncand = cfg->ncand;
    ma.ctuSize = cfg->ctuSize;
    ma.off = s->dOff;
    if (!s->offValid || memcmp(off, s->hOff, sizeof(off)) != 0) {
        if (!stage_upload(s, s->dOff, off, sizeof(off))) return false;
        memcpy(s->hOff, off, sizeof(off));
        s->offValid = true;
    }
    return true;
}

// Enqueues both stages in batched (persistent) form on the session stream.
static bool enqueue_batched(xg_session *s, RefineArgs ra, McArgs ma, bool timed)
{
    const int grid = s->numCUs * s->persistFactor;
    k_zero_counters<<<1, 64, 0, s->stream>>>(s->dCounters, 2);
    ra.counter = s->dCounters;
    ra.taskBegin = 0;
    ra.taskEnd = (int)s->tasks.size();
    int e0 = -1, e1 = -1;
    if (timed) {
        e0 = ev_get(s);
        e1 = ev_get(s);
        if (e0 < 0 || e1 < 0) timed = false;
    }
    if (timed) hipEventRecord(s->evPool[e0], s->stream);
    k_refine<<<std::min(grid, std::max(1, ra.taskEnd)), 64, 0, s->stream>>>(ra);
    if (timed) {
        hipEventRecord(s->evPool[e1], s->stream);
        s->evPending[0].push_back({e0, e1});
    }
    ma.counter = s->dCounters + 1;
    ma.begin = 0;
    ma.end = s->nblocks;
    k_mc<<<std::min(grid, std::max(1, s->nblocks)), 64, 0, s->stream>>>(ma);
    return XG_HIP(hipGetLastError());
}

static bool enqueue_plain(xg_session *s, RefineArgs ra, McArgs ma, bool timed)
{
    int e0 = -1, e1 = -1;
    if (timed) {
        e0 = ev_get(s);
        e1 = ev_get(s);
        if (e0 < 0 || e1 < 0) timed = false;
    }
    if (timed) hipEventRecord(s->evPool[e0], s->stream);
    for (int oi = 0; oi < kNumShapes; oi++) {
        const int b0 = s->shapeTaskBegin[oi], b1 = s->shapeTaskBegin[oi + 1];
        if (b1 <= b0) continue;
        ra.counter = nullptr;
        ra.taskBegin = b0;
        ra.taskEnd = b1;
        k_refine<<<b1 - b0, 64, 0, s->stream>>>(ra);
    }
    if (timed) {
        hipEventRecord(s->evPool[e1], s->stream);
        s->evPending[0].push_back({e0, e1});
    }
    for (int oi = 0; oi < kNumShapes; oi++) {
        const int b0 = s->shapeBlockBegin[oi], b1 = s->shapeBlockBegin[oi + 1];
        if (b1 <= b0) continue;
        ma.counter = nullptr;
        ma.begin = b0;
        ma.end = b1;
        k_mc<<<b1 - b0, 64, 0, s->stream>>>(ma);
    }
    return XG_HIP(hipGetLastError());
}

}  // namespace xg

extern "C" {

int32_t xg_session_create(xg_session **out, int32_t maxBlocks)
{
    if (!out || maxBlocks <= 0) {
        set_error(XG_ERR_ARG, "xg_session_create: bad arguments");
        return XG_ERR_ARG;
    }
    *out = nullptr;
    if (!device_init()) return XG_ERR_HIP;
    xg_session *s = new xg_session();
    s->maxBlocks = maxBlocks;
    for (int i = 0; i < kNumShapes; i++) {
        const int w = 4 << (i / 5), h = 4 << (i % 5);
        // default tile: about 16K luma candidate-samples per task
        s->tile[i] = std::max(1, std::min(XG_MAX_CAND, 16384 / (w * h)));
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        s->numCUs = prop.multiProcessorCount;
    bool ok = XG_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) &&
              XG_HIP(hipMalloc((void **)&s->dPics, sizeof(xg_picture) * kPicSlots)) &&
              XG_HIP(hipMalloc((void **)&s->dBlocks, sizeof(xg_block) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dOrder, sizeof(int) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dCandCost, sizeof(uint32_t) * (size_t)maxBlocks * XG_MAX_CAND)) &&
              XG_HIP(hipMalloc((void **)&s->dBest, sizeof(xg_block_result) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dCounters, 64 * sizeof(int))) &&
              XG_HIP(hipMalloc((void **)&s->dOff, 2 * XG_MAX_CAND * sizeof(int16_t)));
    if (!ok) {
        xg_session_destroy(s);
        return XG_ERR_NOMEM;
    }
    *out = s;
    return XG_OK;
}

int32_t xg_session_destroy(xg_session *s)
{
    if (!s) return XG_OK;
    if (s->stream) hipStreamSynchronize(s->stream);
    fdct_release(s);
    if (s->graphExec) hipGraphExecDestroy(s->graphExec);
    if (s->graph) hipGraphDestroy(s->graph);
    for (auto e : s->evPool) hipEventDestroy(e);
    hipFree(s->dPics);
    hipFree(s->dBlocks);
    hipFree(s->dOrder);
    hipFree(s->dTasks);
    hipFree(s->dCandCost);
    hipFree(s->dBest);
    hipFree(s->dCounters);
    hipFree(s->dOff);
    if (s->hStage) hipHostFree(s->hStage);
    if (s->stream) hipStreamDestroy(s->stream);
    delete s;
    return XG_OK;
}

void *xg_session_stream(xg_session *s) { return s ? (void *)s->stream : nullptr; }

int32_t xg_session_set_picture(xg_session *s, int32_t slot, const xg_picture *pic)
{
    if (!s || !pic || slot < 0 || slot >= kPicSlots) {
        set_error(XG_ERR_ARG, "xg_session_set_picture: bad arguments");
        return XG_ERR_ARG;
    }
    for (int p = 0; p < 3; p++) {
        const xg_plane &pl = pic->plane[p];
        const int sub = p ? 1 : 0;
        if (!pl.data || pl.width != (pic->plane[0].width >> sub) || pl.height != (pic->plane[0].height >> sub) ||
            pl.stride < pl.width + 2 * pl.pad || pl.pad < 0) {
            set_error(XG_ERR_ARG, "xg_session_set_picture: plane %d malformed", p);
            return XG_ERR_ARG;
        }
    }
    s->pics[slot] = *pic;
    s->picSet[slot] = true;
    if (!stage_upload(s, s->dPics + slot, pic, sizeof(xg_picture))) return XG_ERR_HIP;
    return XG_OK;
}

int32_t xg_session_set_blocks(xg_session *s, const xg_block *blocks, int32_t n)
{
    if (!s || (n > 0 && !blocks) || n < 0 || n > s->maxBlocks) {
        set_error(XG_ERR_ARG, "xg_session_set_blocks: bad arguments (n=%d)", n);
        return XG_ERR_ARG;
    }
    for (int i = 0; i < n; i++) {
        const xg_block &B = blocks[i];
        const bool okw = B.w >= 4 && B.w <= 64 && !(B.w & (B.w - 1));
        const bool okh = B.h >= 4 && B.h <= 64 && !(B.h & (B.h - 1));
        if (!okw || !okh) {
            set_error(XG_ERR_ARG, "xg_session_set_blocks: block %d has size %dx%d", i, B.w, B.h);
            return XG_ERR_ARG;
        }
    }
    s->blocks.assign(blocks, blocks + n);
    s->nblocks = n;
    s->blocksVersion++;
    if (n && !stage_upload(s, s->dBlocks, blocks, sizeof(xg_block) * (size_t)n)) return XG_ERR_HIP;
    return XG_OK;
}

int32_t xg_session_set_tile(xg_session *s, int32_t lgW, int32_t lgH, int32_t candPerTask)
{
    if (!s || lgW < 2 || lgW > 6 || lgH < 2 || lgH > 6 || candPerTask < 1 || candPerTask > XG_MAX_CAND) {
        set_error(XG_ERR_ARG, "xg_session_set_tile: bad arguments");
        return XG_ERR_ARG;
    }
    s->tile[(lgW - 2) * 5 + (lgH - 2)] = candPerTask;
    return XG_OK;
}

int32_t xg_session_get_tile(xg_session *s, int32_t lgW, int32_t lgH)
{
    if (!s || lgW < 2 || lgW > 6 || lgH < 2 || lgH > 6) return -1;
    return s->tile[(lgW - 2) * 5 + (lgH - 2)];
}

int32_t xg_refine_frac(xg_session *s, const xg_refine_cfg *cfg, const xg_picture *pred, int32_t mode)
{
    if (!s || !validate_cfg(cfg) || !pred) {
        if (s && !pred) set_error(XG_ERR_ARG, "xg_refine_frac: prediction picture required");
        return XG_ERR_ARG;
    }
    if (s->nblocks == 0) return XG_OK;
    if (!validate_blocks(s, pred, cfg->ctuSize)) return XG_ERR_ARG;
    if (!build_tasks(s, cfg->ncand)) return XG_ERR_HIP;
    RefineArgs ra;
    McArgs ma;
    if (!fill_args(s, cfg, ra, ma, pred)) return XG_ERR_HIP;
    s->lastNcand = cfg->ncand;
    if (mode == XG_MODE_PLAIN) return enqueue_plain(s, ra, ma, s->timing) ? XG_OK : XG_ERR_HIP;
    if (mode == XG_MODE_BATCHED) return enqueue_batched(s, ra, ma, s->timing) ? XG_OK : XG_ERR_HIP;
    if (mode != XG_MODE_GRAPH) {
        set_error(XG_ERR_ARG, "xg_refine_frac: unknown mode %d", mode);
        return XG_ERR_ARG;
    }
    // graph key: everything captured by value in the kernel arguments
    uint64_t key = mix(s->tasksKey, (uint64_t)(uintptr_t)pred->plane[0].data);
    key = mix(key, (uint64_t)(uintptr_t)pred->plane[1].data);
    key = mix(key, (uint64_t)(uintptr_t)pred->plane[2].data);
    key = mix(key, ((uint64_t)cfg->lambda << 32) | (uint32_t)cfg->ctuSize);
    key = mix(key, (uint64_t)cfg->useChroma);
    for (int i = 0; i < cfg->ncand; i++) key = mix(key, ((uint64_t)(uint16_t)cfg->offx[i] << 16) | (uint16_t)cfg->offy[i]);
    if (!s->graphExec || key != s->graphKey) {
        if (s->graphExec) hipGraphExecDestroy(s->graphExec);
        if (s->graph) hipGraphDestroy(s->graph);
        s->graphExec = nullptr;
        s->graph = nullptr;
        if (!XG_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal))) return XG_ERR_HIP;
        const bool ok = enqueue_batched(s, ra, ma, false);
        hipGraph_t g = nullptr;
        const bool ended = XG_HIP(hipStreamEndCapture(s->stream, &g));
        if (!ok || !ended) return XG_ERR_HIP;
        s->graph = g;
        if (!XG_HIP(hipGraphInstantiate(&s->graphExec, g, nullptr, nullptr, 0))) return XG_ERR_HIP;
        s->graphKey = key;
    }
    int e0 = -1, e1 = -1;
    bool timed = s->timing;
    if (timed) {
        e0 = ev_get(s);
        e1 = ev_get(s);
        if (e0 < 0 || e1 < 0) timed = false;
    }
    if (timed) hipEventRecord(s->evPool[e0], s->stream);
    if (!XG_HIP(hipGraphLaunch(s->graphExec, s->stream))) return XG_ERR_HIP;
    if (timed) {
        hipEventRecord(s->evPool[e1], s->stream);
        s->evPending[1].push_back({e0, e1});
    }
    return XG_OK;
}

int32_t xg_session_outputs(xg_session *s, uint32_t **candCost, xg_block_result **best)
{
    if (!s) return XG_ERR_ARG;
    if (candCost) *candCost = s->dCandCost;
    if (best) *best = s->dBest;
    return XG_OK;
}

int32_t xg_session_sync(xg_session *s)
{
    if (!s) return XG_ERR_ARG;
    if (!XG_HIP(hipStreamSynchronize(s->stream))) return XG_ERR_HIP;
    if (s->timing) ev_collect(s);
    return XG_OK;
}

int32_t xg_session_read(xg_session *s, uint32_t *candCost, xg_block_result *best)
{
    if (!s) return XG_ERR_ARG;
    if (xg_session_sync(s) != XG_OK) return XG_ERR_HIP;
    if (candCost && s->nblocks && s->lastNcand &&
        !XG_HIP(hipMemcpy(candCost, s->dCandCost, sizeof(uint32_t) * (size_t)s->nblocks * s->lastNcand,
                          hipMemcpyDeviceToHost)))
        return XG_ERR_HIP;
    if (best && s->nblocks &&
        !XG_HIP(hipMemcpy(best, s->dBest, sizeof(xg_block_result) * (size_t)s->nblocks, hipMemcpyDeviceToHost)))
        return XG_ERR_HIP;
    return XG_OK;
}

int32_t xg_session_set_timing(xg_session *s, int32_t enable)
{
    if (!s) return XG_ERR_ARG;
    if (s->timing && !enable) xg_session_sync(s);
    s->timing = enable != 0;
    return XG_OK;
}

int32_t xg_session_reset_timers(xg_session *s)
{
    if (!s) return XG_ERR_ARG;
    xg_session_sync(s);
    s->accMs[0] = s->accMs[1] = 0;
    s->accN[0] = s->accN[1] = 0;
    return XG_OK;
}

// stage 0: refinement kernel(s) of the eager modes; stage 1: whole graph replay.
int32_t xg_session_kernel_time(xg_session *s, int32_t stage, float *avgMs, int32_t *count)
{
    if (!s || stage < 0 || stage > 1) return XG_ERR_ARG;
    if (xg_session_sync(s) != XG_OK) return XG_ERR_HIP;
    if (avgMs) *avgMs = s->accN[stage] ? (float)(s->accMs[stage] / s->accN[stage]) : 0.f;
    if (count) *count = s->accN[stage];
    return XG_OK;
}

int32_t xg_session_autotune(xg_session *s, const xg_refine_cfg *cfg, int32_t reps, const char *path)
{
    if (!s || !validate_cfg(cfg) || reps < 1) return XG_ERR_ARG;
    if (s->nblocks == 0) return XG_OK;
    // Time each shape's refinement launch alone for every legal tile.
    std::vector<xg_block> all = s->blocks;
    hipEvent_t e0, e1;
    if (!XG_HIP(hipEventCreate(&e0)) || !XG_HIP(hipEventCreate(&e1))) return XG_ERR_HIP;
    int32_t rc = XG_OK;
    for (int sh = 0; sh < kNumShapes && rc == XG_OK; sh++) {
        const int w = 4 << (sh / 5), h = 4 << (sh % 5);
        std::vector<xg_block> sub;
        for (auto &B : all)
            if (B.w == w && B.h == h) sub.push_back(B);
        if (sub.empty()) continue;
        if (xg_session_set_blocks(s, sub.data(), (int)sub.size()) != XG_OK) { rc = XG_ERR_HIP; break; }
        if (!validate_blocks(s, nullptr, cfg->ctuSize)) { rc = XG_ERR_ARG; break; }
        float bestMs = 1e30f;
        int bestT = s->tile[sh];
        for (int cpt = 1; cpt <= cfg->ncand; cpt <<= 1) {
            s->tile[sh] = cpt;
            if (!build_tasks(s, cfg->ncand)) { rc = XG_ERR_HIP; break; }
            RefineArgs ra;
            McArgs ma;
            if (!fill_args(s, cfg, ra, ma, nullptr)) { rc = XG_ERR_HIP; break; }
            ra.counter = s->dCounters;
            ra.taskBegin = 0;
            ra.taskEnd = (int)s->tasks.size();
            const int grid = std::min(s->numCUs * s->persistFactor, std::max(1, ra.taskEnd));
            float tot = 0;
            for (int r = 0; r <= reps; r++) {
                k_zero_counters<<<1, 64, 0, s->stream>>>(s->dCounters, 2);
                hipEventRecord(e0, s->stream);
                k_refine<<<grid, 64, 0, s->stream>>>(ra);
                hipEventRecord(e1, s->stream);
                if (!XG_HIP(hipEventSynchronize(e1))) { rc = XG_ERR_HIP; break; }
                float ms = 0;
                hipEventElapsedTime(&ms, e0, e1);
                if (r > 0) tot += ms;    // first run warms up
            }
            if (rc != XG_OK) break;
            if (tot < bestMs) { bestMs = tot; bestT = cpt; }
        }
        s->tile[sh] = bestT;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (xg_session_set_blocks(s, all.data(), (int)all.size()) != XG_OK && rc == XG_OK) rc = XG_ERR_HIP;
    if (rc == XG_OK && path) {
        FILE *f = fopen(path, "w");
        if (!f) {
            set_error(XG_ERR_ARG, "xg_session_autotune: cannot write %s", path);
            return XG_ERR_ARG;
        }
        fprintf(f, "# xg266 refine tiles: lgW lgH candidates_per_task (ncand %d)\n", cfg->ncand);
        for (int sh = 0; sh < kNumShapes; sh++) fprintf(f, "%d %d %d\n", sh / 5 + 2, sh % 5 + 2, s->tile[sh]);
        fclose(f);
    }
    return rc;
}

int32_t xg_session_load_tiles(xg_session *s, const char *path)
{
    if (!s || !path) return XG_ERR_ARG;
    FILE *f = fopen(path, "r");
    if (!f) {
        set_error(XG_ERR_ARG, "xg_session_load_tiles: cannot read %s", path);
        return XG_ERR_ARG;
    }
    char line[256];
    int n = 0;
    while (fgets(line, sizeof(line), f)) {
        int lw, lh, t;
        if (line[0] == '#') continue;
        if (sscanf(line, "%d %d %d", &lw, &lh, &t) == 3 && xg_session_set_tile(s, lw, lh, t) == XG_OK) n++;
    }
    fclose(f);
    return n > 0 ? XG_OK : XG_ERR_ARG;
}

}  // extern "C"

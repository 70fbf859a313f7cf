// xg_pipeline.hip -- the full motion-search pipeline of a frame with two
// reference pictures, on a raster grid of 16x16 blocks:
//
//   1. k_pipe_full: integer full search +/-range around (0,0) per block and
//      reference (the closed Xin26xMotionEstFull, h26x_core/include/h26x_motion_est.h:26:
//      window intersected with the CU MV window of h266_enc_init.c:1462-1465,
//      cost (SAD << 8) + bits(mvd) * sadLambda, raster order, earliest wins);
//   2. k_pipe_dia: predictor = component-wise median of the left, top and
//      top-right blocks' full-search MVs; start candidates as
//      Xin266FillMeCand (h266_analyze_inter_cu.c:133-293) from the predictor
//      and the own / left / top / top-right MVs; diamond search (the closed
//      Xin26xMotionEstDia); then the half- and quarter-pel rings of
//      Xin266FullSubPelMe (h266_sub_pel_refine.c:81-212) on 8-tap
//      interpolated SAD;
//   3. k_pipe_select: per block the reference with the lower quarter-pel cost
//      (ties: reference 0), written into the session's block list;
//   4. the session's persistent refinement kernel: 1/16-pel candidates around
//      the quarter-pel MV (SATD + chroma) and the motion compensation of the
//      winner.
// All four run on the session stream; in graph mode they are one captured
// HIP graph per frame shape.
//
// k_pipe_full is the dominant kernel (one 4-wave workgroup per block and
// reference).  The window, (16 + 2 range)^2 samples, is staged once in LDS;
// each lane costs four horizontally adjacent candidates at once from ten
// dword reads per row (16 + 3 samples): candidates dx and dx+2 take the
// sample pairs as read, dx+1 and dx+3 the pairs realigned by v_alignbit, and
// v_sad_u16 sums a pair's two absolute differences, so a row costs 9
// realignments and 32 v_sad_u16 for 4 x 16 absolute differences.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string.h>
#include <vector>

#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

constexpr int kPB = 16;                          // block side of the pipeline grid
constexpr int kPipeMaxRange = 64;
constexpr int kPipeQuads = 3;                              // candidate quads per lane in the full search
// window row: 2 * range + 16 samples plus the over-read of the last lane's quads
constexpr int kPipeWinW = 2 * kPipeMaxRange + kPB + 4 * kPipeQuads + 2;   // 158
constexpr int kPipeWinH = 2 * kPipeMaxRange + kPB + 1;    // 145: a two-row item may read one row past the window

struct PipeArgs {
    const xg_picture *pics;
    int curSlot, refSlot0, refSlot1;
    int fw, fh, nbx, nb, range, ctu;
    uint32_t lambda;
    int16_t *fullMv, *predMv, *intMv, *qMv;    // [nb][2 refs][2]
    uint32_t *intCost, *qCost;                 // [nb][2 refs]
    xg_block *blocks;                          // session block list (refinement input)
};

struct IWin { int mnx, mny, mxx, mxy; };       // CU MV window in integer pel

__device__ __forceinline__ IWin int_window(int x, int y, int fw, int fh, int ctu)
{
    IWin w;
    w.mnx = (int16_t)(-x - ctu - 8);
    w.mny = (int16_t)(-y - ctu - 8);
    w.mxx = (int16_t)(fw + (ctu - kPB + 8) - x);
    w.mxy = (int16_t)(fh + (ctu - kPB + 8) - y);
    return w;
}

// v_sad_u16 on packed pairs: |a.lo - b.lo| + |a.hi - b.hi| + acc (checked on
// gfx950: both 16-bit halves take part).
__device__ __forceinline__ uint32_t sad2(uint32_t a, uint32_t b, uint32_t acc)
{
    return __builtin_amdgcn_sad_u16(a, b, acc);
}

__device__ __forceinline__ unsigned long long min_u64(unsigned long long a, unsigned long long b)
{
    return a < b ? a : b;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long k)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) k = min_u64(k, (unsigned long long)__shfl_xor((long long)k, m, 64));
    return k;
}

// ---------------------------------------------------------------------------
// 1. Full search
// ---------------------------------------------------------------------------
// Each lane costs NQ quads (4 * NQ horizontally adjacent candidates) of one
// search row: the window row's dwords D and their one-sample realignment F are
// shared by all of them, so the loads and v_alignbit per v_sad_u16 fall as NQ
// grows.
template <int NQ>
__global__ void __launch_bounds__(256) k_pipe_full(PipeArgs a)
{
    __shared__ __align__(16) uint16_t win[kPipeWinH * kPipeWinW];
    __shared__ __align__(16) uint32_t curL[kPB * kPB / 2];     // packed pairs
    __shared__ unsigned long long red[4];
    const int b = blockIdx.x >> 1, r = blockIdx.x & 1;
    const int x = (b % a.nbx) * kPB, y = (b / a.nbx) * kPB;
    const IWin w = int_window(x, y, a.fw, a.fh, a.ctu);
    const int x0 = max(-a.range, w.mnx), x1 = min(a.range, w.mxx);
    const int y0 = max(-a.range, w.mny), y1 = min(a.range, w.mxy);
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1;
    const xg_plane &RP = a.pics[r ? a.refSlot1 : a.refSlot0].plane[0];
    const xg_plane &CP = a.pics[a.curSlot].plane[0];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    {   // window rows (ny + 15) x (nx + 15) samples: a wave per row, lanes along it,
        // 4 rows x 3 columns of loads in flight per lane
        const int rows = ny + kPB - 1, cols = nx + kPB - 1;
        const uint16_t *src = RP.data + (intptr_t)(y + y0) * RP.stride + (x + x0);
        for (int r0 = wave; r0 < rows; r0 += 16) {
            uint16_t v[4][3];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint16_t *sp = src + (intptr_t)min(r0 + 4 * k, rows - 1) * RP.stride;
#pragma unroll
                for (int c = 0; c < 3; c++) v[k][c] = sp[min(lane + 64 * c, cols - 1)];
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (r0 + 4 * k < rows)
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        if (lane + 64 * c < cols) win[(r0 + 4 * k) * kPipeWinW + lane + 64 * c] = v[k][c];
        }
        if (tid < kPB * kPB / 2)
            curL[tid] = *reinterpret_cast<const uint32_t *>(CP.data + (intptr_t)(y + (tid >> 3)) * CP.stride + x +
                                                            2 * (tid & 7));
    }
    __syncthreads();
    constexpr int ND = 2 * NQ + 8;                  // window dwords per row read by one lane
    constexpr int NK = 4 * NQ;
    // an item: NK adjacent candidates of two search rows; window row i serves current row i of the
    // first and current row i - 1 of the second, so each window row is read and realigned once
    const int ng = (nx + NK - 1) / NK, items = ((ny + 1) >> 1) * ng;
    const uint32_t *w32 = reinterpret_cast<const uint32_t *>(win);
    unsigned long long best = ~0ull;
    for (int it = tid; it < items; it += 256) {
        const int p = it / ng, g = it - p * ng, dyi = 2 * p;
        const uint32_t *rp = w32 + ((dyi * kPipeWinW) >> 1) + 2 * NQ * g;
        uint32_t acc0[NK], acc1[NK];
#pragma unroll
        for (int k = 0; k < NK; k++) acc0[k] = acc1[k] = 0;
        uint32_t D[ND], F[ND - 1], C[8], Cp[8];
        auto load_row = [&](int i) {
#pragma unroll
            for (int j = 0; j < ND; j++) D[j] = rp[i * (kPipeWinW / 2) + j];
#pragma unroll
            for (int j = 0; j < ND - 1; j++) F[j] = align16(D[j + 1], D[j], 16);    // pairs starting at odd samples
        };
        auto load_cur = [&](int i) {
            const uint4 *c4 = reinterpret_cast<const uint4 *>(curL + i * (kPB / 2));
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const uint4 v = c4[j];
                C[4 * j] = v.x;
                C[4 * j + 1] = v.y;
                C[4 * j + 2] = v.z;
                C[4 * j + 3] = v.w;
            }
        };
        auto sads = [&](const uint32_t *Cr, uint32_t *acc) {
#pragma unroll
            for (int t = 0; t < NQ; t++)
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    acc[4 * t] = sad2(D[2 * t + j], Cr[j], acc[4 * t]);
                    acc[4 * t + 1] = sad2(F[2 * t + j], Cr[j], acc[4 * t + 1]);
                    acc[4 * t + 2] = sad2(D[2 * t + j + 1], Cr[j], acc[4 * t + 2]);
                    acc[4 * t + 3] = sad2(F[2 * t + j + 1], Cr[j], acc[4 * t + 3]);
                }
        };
        load_row(0);
        load_cur(0);
        sads(C, acc0);
#pragma unroll 2
        for (int i = 1; i < kPB; i++) {
#pragma unroll
            for (int j = 0; j < 8; j++) Cp[j] = C[j];
            load_row(i);
            load_cur(i);
            sads(C, acc0);
            sads(Cp, acc1);
        }
        load_row(kPB);
        sads(C, acc1);
        // the lane's candidates come in raster order, so a strict < keeps the lowest index
        // among equal costs, as the 64-bit key order does
        uint32_t bc = 0xFFFFFFFFu, bi = 0;
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int dy = dyi + r;
            if (dy < ny) {
                const uint32_t by = eg_bits(4 * (y0 + dy));
#pragma unroll
                for (int k = 0; k < NK; k++) {
                    const int ox = NK * g + k;
                    const uint32_t cost = ((r ? acc1[k] : acc0[k]) << 8) + (eg_bits(4 * (x0 + ox)) + by) * a.lambda;
                    if (x0 + ox <= x1 && cost < bc) {
                        bc = cost;
                        bi = (uint32_t)(dy * nx + ox);
                    }
                }
            }
        }
        if (bc != 0xFFFFFFFFu) best = min_u64(best, ((unsigned long long)bc << 32) | bi);
    }
    best = wave_min_u64(best);
    if (lane == 0) red[wave] = best;
    __syncthreads();
    if (tid == 0) {
        best = min_u64(min_u64(red[0], red[1]), min_u64(red[2], red[3]));
        const int idx = (int)(uint32_t)best;
        a.fullMv[(b * 2 + r) * 2] = (int16_t)(x0 + idx % nx);
        a.fullMv[(b * 2 + r) * 2 + 1] = (int16_t)(y0 + idx / nx);
    }
}

// ---------------------------------------------------------------------------
// 2. Predictor, candidates, diamond, half/quarter-pel rings (one wave per
//    block and reference)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int median3(int p, int q, int s)
{
    const int lo = min(p, q), hi = max(p, q);
    return min(max(s, lo), hi);
}

// SAD of the current block's row `row` (packed pairs c[0..7]) against the
// reference row at sample (rx, ry) of plane R (16 samples, any parity).
__device__ __forceinline__ uint32_t row_sad(const xg_plane &R, int rx, int ry, const uint32_t *c)
{
    const uint16_t *p = R.data + (intptr_t)ry * R.stride + rx;
    const uint32_t *q = reinterpret_cast<const uint32_t *>(reinterpret_cast<uintptr_t>(p) & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(reinterpret_cast<uintptr_t>(p) & 2) << 3;
    uint32_t D[9];
#pragma unroll
    for (int j = 0; j < 9; j++) D[j] = q[j];
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) acc = sad2(align16(D[j + 1], D[j], sh), c[j], acc);
    return acc;
}

// Sum over the 16 lanes of a row of the wavefront.
__device__ __forceinline__ uint32_t sum16(uint32_t v)
{
    v += (uint32_t)shfl_xor<1>((int)v);
    v += (uint32_t)shfl_xor<2>((int)v);
    v += (uint32_t)shfl_xor<4>((int)v);
    v += (uint32_t)shfl_xor<8>((int)v);
    return v;
}

// Sum over aligned groups of 8 lanes.
__device__ __forceinline__ uint32_t sum8(uint32_t v)
{
    v += (uint32_t)shfl_xor<1>((int)v);
    v += (uint32_t)shfl_xor<2>((int)v);
    v += (uint32_t)shfl_xor<4>((int)v);
    return v;
}

__device__ __forceinline__ uint32_t cost_q(uint32_t sad, int mvqx, int mvqy, int px, int py, uint32_t lambda)
{
    return (sad << 8) + (eg_bits(mvqx - px) + eg_bits(mvqy - py)) * lambda;
}

__global__ void __launch_bounds__(64) k_pipe_dia(PipeArgs a)
{
    const int b = blockIdx.x >> 1, r = blockIdx.x & 1;
    const int bx = b % a.nbx, by = b / a.nbx, x = bx * kPB, y = by * kPB;
    const IWin w = int_window(x, y, a.fw, a.fh, a.ctu);
    const xg_plane &RP = a.pics[r ? a.refSlot1 : a.refSlot0].plane[0];
    const xg_plane &CP = a.pics[a.curSlot].plane[0];
    const int lane = threadIdx.x, grp = lane >> 4, row = lane & 15;

    // predictor and start candidates (every lane computes the same list)
    int nqx[4], nqy[4];
    {
        const int nbi[4] = {b, bx > 0 ? b - 1 : -1, by > 0 ? b - a.nbx : -1,
                            (by > 0 && bx + 1 < a.nbx) ? b - a.nbx + 1 : -1};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            nqx[k] = nbi[k] < 0 ? 0 : 4 * a.fullMv[(nbi[k] * 2 + r) * 2];
            nqy[k] = nbi[k] < 0 ? 0 : 4 * a.fullMv[(nbi[k] * 2 + r) * 2 + 1];
        }
    }
    const int px = (int16_t)median3(nqx[1], nqx[2], nqx[3]), py = (int16_t)median3(nqy[1], nqy[2], nqy[3]);
    int cx[6], cy[6], nc = 1;
    cx[0] = 0;
    cy[0] = 0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int vx = i == 0 ? px : nqx[i - 1], vy = i == 0 ? py : nqy[i - 1];
        const int ix = (int16_t)((vx + 2) >> 2), iy = (int16_t)((vy + 2) >> 2);
        bool dup = false;
#pragma unroll
        for (int j = 0; j < 6; j++) dup |= (j < nc) && cx[j] == ix && cy[j] == iy;
        if (!dup) {
#pragma unroll
            for (int j = 0; j < 6; j++)
                if (j == nc) {
                    cx[j] = ix;
                    cy[j] = iy;
                }
            nc++;
        }
    }
    // the lane's current row as packed pairs
    uint32_t c[8];
    {
        const uint32_t *cp = reinterpret_cast<const uint32_t *>(CP.data + (intptr_t)(y + row) * CP.stride + x);
#pragma unroll
        for (int j = 0; j < 8; j++) c[j] = cp[j];
    }
    // cost of up to 4 integer positions at once (group g = lane / 16 takes position g)
    auto eval4 = [&](const int *mx, const int *my, const bool *ok, uint32_t *cost, uint32_t *sads) {
        const int g = grp;
        uint32_t s = 0;
        if (ok[0] | ok[1] | ok[2] | ok[3]) {
            const int gx = g == 0 ? mx[0] : g == 1 ? mx[1] : g == 2 ? mx[2] : mx[3];
            const int gy = g == 0 ? my[0] : g == 1 ? my[1] : g == 2 ? my[2] : my[3];
            const bool gok = g == 0 ? ok[0] : g == 1 ? ok[1] : g == 2 ? ok[2] : ok[3];
            if (gok) s = row_sad(RP, x + gx, y + gy + row, c);
            s = sum16(s);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            sads[k] = (uint32_t)__shfl((int)s, 16 * k, 64);
            cost[k] = ok[k] ? cost_q(sads[k], 4 * mx[k], 4 * my[k], px, py, a.lambda) : 0xFFFFFFFFu;
        }
    };
    // start candidates in list order (clipped to the MV window), strict <
    uint32_t best = 0xFFFFFFFFu;
    int bmx = 0, bmy = 0;
    for (int base = 0; base < nc; base += 4) {
        int mx[4], my[4];
        bool ok[4];
        uint32_t cost[4], sads[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int ix = 0, iy = 0;
#pragma unroll
            for (int j = 0; j < 6; j++)
                if (j == base + k) {
                    ix = cx[j];
                    iy = cy[j];
                }
            mx[k] = min(max(ix, w.mnx), w.mxx);
            my[k] = min(max(iy, w.mny), w.mxy);
            ok[k] = base + k < nc;
        }
        eval4(mx, my, ok, cost, sads);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (cost[k] < best) {
                best = cost[k];
                bmx = mx[k];
                bmy = my[k];
            }
    }
    // small diamond L, R, T, B around the best, inside the window and +/-range of the rounded predictor
    const int ccx = (px + 2) >> 2, ccy = (py + 2) >> 2;
    for (int step = 0; step < 2 * a.range; step++) {
        const int dxs[4] = {-1, 1, 0, 0}, dys[4] = {0, 0, -1, 1};
        int mx[4], my[4];
        bool ok[4];
        uint32_t cost[4], sads[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            mx[k] = bmx + dxs[k];
            my[k] = bmy + dys[k];
            ok[k] = mx[k] >= w.mnx && mx[k] <= w.mxx && my[k] >= w.mny && my[k] <= w.mxy &&
                    mx[k] >= ccx - a.range && mx[k] <= ccx + a.range && my[k] >= ccy - a.range &&
                    my[k] <= ccy + a.range;
        }
        eval4(mx, my, ok, cost, sads);
        uint32_t nb = best;
        int nx = bmx, ny = bmy;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (cost[k] < nb) {
                nb = cost[k];
                nx = mx[k];
                ny = my[k];
            }
        if (nx == bmx && ny == bmy) break;
        bmx = nx;
        bmy = ny;
        best = nb;
    }
    if (lane == 0) {
        a.predMv[(b * 2 + r) * 2] = (int16_t)px;
        a.predMv[(b * 2 + r) * 2 + 1] = (int16_t)py;
        a.intMv[(b * 2 + r) * 2] = (int16_t)bmx;
        a.intMv[(b * 2 + r) * 2 + 1] = (int16_t)bmy;
        a.intCost[b * 2 + r] = best;
    }

    // half- then quarter-pel rings (L, R, T, B, TL, TR, BL, BR), 8-tap SAD, strict <;
    // lane = candidate * 8 + m: columns 2m, 2m+1 of candidate lane / 8
    int qx = bmx * 4, qy = bmy * 4;
    if (bmx < w.mnx || bmx > w.mxx || bmy < w.mny || bmy > w.mxy) {   // cannot happen: kept for the reference's guard
        if (lane == 0) {
            a.qMv[(b * 2 + r) * 2] = (int16_t)qx;
            a.qMv[(b * 2 + r) * 2 + 1] = (int16_t)qy;
            a.qCost[b * 2 + r] = best;
        }
        return;
    }
    const int cand = lane >> 3, m = lane & 7;
    const int ringx[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, ringy[8] = {0, 0, -1, 1, -1, -1, 1, 1};
    uint32_t cw[16];
    {
        const uint32_t *cp = reinterpret_cast<const uint32_t *>(CP.data + (intptr_t)y * CP.stride + x + 2 * m);
#pragma unroll
        for (int i = 0; i < 16; i++) cw[i] = cp[(intptr_t)i * (CP.stride >> 1)];
    }
    for (int pass = 0; pass < 2; pass++) {
        const int step = pass == 0 ? 2 : 1;
        int rx = ringx[0], ry = ringy[0];
#pragma unroll
        for (int k = 1; k < 8; k++)
            if (k == cand) {
                rx = ringx[k];
                ry = ringy[k];
            }
        const int tx = qx + rx * step, ty = qy + ry * step;          // quarter pel
        const int m16x = tx * 4, m16y = ty * 4;
        const int fx = m16x & 15, fy = m16y & 15;
        uint32_t hp[4];
        int vc[8];
        {
            const int8_t *ch = taps_for<8>(fx, 0), *cv = taps_for<8>(fy, 0);
#pragma unroll
            for (int j = 0; j < 4; j++)
                hp[j] = pack16(ch[2 * j], ch[2 * j + 1]);
#pragma unroll
            for (int j = 0; j < 8; j++) vc[j] = cv[j];
        }
        // first tap of column 2m at row 0: (x + 2m + (m16x >> 4) - 3, y + (m16y >> 4) - 3)
        const int sx = x + 2 * m + (m16x >> 4) - 3, sy = y + (m16y >> 4) - 3;
        int t0[kPB + 7], t1[kPB + 7];
        uint32_t sad = 0;
#pragma unroll
        for (int i = 0; i < kPB + 7; i++) {
            const uint16_t *p = RP.data + (intptr_t)(sy + i) * RP.stride + sx;
            const uint32_t *q = reinterpret_cast<const uint32_t *>(reinterpret_cast<uintptr_t>(p) & ~(uintptr_t)3);
            const uint32_t sh = (uint32_t)(reinterpret_cast<uintptr_t>(p) & 2) << 3;
            uint32_t D[6], E[5];
#pragma unroll
            for (int j = 0; j < 6; j++) D[j] = q[j];
#pragma unroll
            for (int j = 0; j < 5; j++) E[j] = align16(D[j + 1], D[j], sh);
            int s0 = kHOffset, s1 = kHOffset;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                s0 = dot2(E[j], hp[j], s0);
                s1 = dot2(align16(E[j + 1], E[j], 16), hp[j], s1);
            }
            t0[i] = s0 >> kHShift;
            t1[i] = s1 >> kHShift;
            if (i >= 7) {
                int v0 = kVOffset, v1 = kVOffset;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    v0 += vc[j] * t0[i - 7 + j];
                    v1 += vc[j] * t1[i - 7 + j];
                }
                sad = sad2(pack16(clip_pix(v0 >> kVShift), clip_pix(v1 >> kVShift)), cw[i - 7], sad);
            }
        }
        sad = sum8(sad);
        const uint32_t cost = cost_q(sad, tx, ty, px, py, a.lambda);
        // earliest candidate among the strictly better ones
        unsigned long long key = ((unsigned long long)cost << 32) | (uint32_t)cand;
#pragma unroll
        for (int mm = 8; mm < 64; mm <<= 1) key = min_u64(key, (unsigned long long)__shfl_xor((long long)key, mm, 64));
        if ((uint32_t)(key >> 32) < best) {
            best = (uint32_t)(key >> 32);
            const int kk = (int)(uint32_t)key;
            qx = qx + ringx[kk] * step;
            qy = qy + ringy[kk] * step;
        }
    }
    if (lane == 0) {
        a.qMv[(b * 2 + r) * 2] = (int16_t)qx;
        a.qMv[(b * 2 + r) * 2 + 1] = (int16_t)qy;
        a.qCost[b * 2 + r] = best;
    }
}

// ---------------------------------------------------------------------------
// 3. Reference choice -> refinement block list
// ---------------------------------------------------------------------------
__global__ void k_pipe_select(PipeArgs a)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.nb) return;
    const int r = a.qCost[b * 2 + 1] < a.qCost[b * 2] ? 1 : 0;
    xg_block B;
    B.x = (int16_t)((b % a.nbx) * kPB);
    B.y = (int16_t)((b / a.nbx) * kPB);
    B.w = kPB;
    B.h = kPB;
    B.pic = (uint8_t)a.curSlot;
    B.ref = (uint8_t)(r ? a.refSlot1 : a.refSlot0);
    B.mvx = 4 * a.qMv[(b * 2 + r) * 2];
    B.mvy = 4 * a.qMv[(b * 2 + r) * 2 + 1];
    B.mvpx = 4 * a.predMv[(b * 2 + r) * 2];
    B.mvpy = 4 * a.predMv[(b * 2 + r) * 2 + 1];
    a.blocks[b] = B;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct PipeState {
    int fw = 0, fh = 0, nb = 0, cap = 0;
    uint64_t gridVersion = 0;          // session blocksVersion right after the grid was installed
    int16_t *dMv = nullptr;            // fullMv | predMv | intMv | qMv, 4 x nb x 4 int16
    uint32_t *dCost = nullptr;         // intCost | qCost, 2 x nb x 2
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t key = 0;
    PipeArgs args;
};

void pipe_release(xg_session *s)
{
    PipeState *p = static_cast<PipeState *>(session_pipe(s));
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->dMv);
    (void)hipFree(p->dCost);
    delete p;
    session_pipe(s) = nullptr;
}

static uint64_t hmix(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h;
}

static bool enqueue_pipe(const PipeState *p, hipStream_t st, xg_session *s, bool timed)
{
    int ev[2];
    const bool t3 = timed && session_timer_begin(s, 3, ev);
    k_pipe_full<kPipeQuads><<<2 * p->nb, 256, 0, st>>>(p->args);
    if (t3) session_timer_end(s, 3, ev);
    k_pipe_dia<<<2 * p->nb, 64, 0, st>>>(p->args);
    k_pipe_select<<<(p->nb + 255) / 256, 256, 0, st>>>(p->args);
    return XG_HIP(hipGetLastError()) && refine_enqueue_prepared(s);
}

}  // namespace xg

using namespace xg;

extern "C" {

int32_t xg_me_pipeline(xg_session *s, const xg_pipeline_cfg *cfg, const xg_picture *cur, const xg_picture *ref0,
                       const xg_picture *ref1, const xg_picture *pred, int32_t mode)
{
    if (!s || !cfg || !cur || !ref0 || !ref1 || !pred || cfg->range < 0 || cfg->range > kPipeMaxRange ||
        cfg->ctuSize < 16 || cfg->ctuSize > 128 || (mode != XG_MODE_BATCHED && mode != XG_MODE_GRAPH)) {
        set_error(XG_ERR_ARG, "xg_me_pipeline: bad arguments");
        return XG_ERR_ARG;
    }
    const int fw = cur->plane[0].width, fh = cur->plane[0].height;
    if (fw < kPB || fh < kPB || (fw % kPB) || (fh % kPB)) {
        set_error(XG_ERR_ARG, "xg_me_pipeline: the frame must be a whole number of 16x16 blocks");
        return XG_ERR_ARG;
    }
    for (const xg_picture *r : {ref0, ref1})
        if (r->plane[0].width != fw || r->plane[0].height != fh || r->plane[0].pad < cfg->ctuSize + 16 ||
            r->plane[1].pad < cfg->ctuSize / 2 + 8) {
            set_error(XG_ERR_ARG, "xg_me_pipeline: reference size/padding (luma pad >= ctuSize + 16)");
            return XG_ERR_ARG;
        }
    const int nbx = fw / kPB, nb = nbx * (fh / kPB);
    SessionView v;
    session_view(s, &v);
    PipeState *p = static_cast<PipeState *>(session_pipe(s));
    if (!p) {
        p = new PipeState();
        session_pipe(s) = p;
    }
    int32_t rc;
    if ((rc = xg_session_set_picture(s, 0, cur)) != XG_OK || (rc = xg_session_set_picture(s, 1, ref0)) != XG_OK ||
        (rc = xg_session_set_picture(s, 2, ref1)) != XG_OK)
        return rc;
    // (re)install the grid when the frame changed or the caller replaced the session's blocks
    if (p->nb != nb || p->fw != fw || p->fh != fh || p->gridVersion != v.blocksVersion) {
        std::vector<xg_block> grid(nb);
        for (int b = 0; b < nb; b++) {
            xg_block &B = grid[b];
            memset(&B, 0, sizeof(B));
            B.x = (int16_t)((b % nbx) * kPB);
            B.y = (int16_t)((b / nbx) * kPB);
            B.w = kPB;
            B.h = kPB;
            B.pic = 0;
            B.ref = 1;
        }
        if ((rc = xg_session_set_blocks(s, grid.data(), nb)) != XG_OK) return rc;
        if (nb > p->cap) {
            if (!XG_HIP(hipStreamSynchronize(v.stream))) return XG_ERR_HIP;
            (void)hipFree(p->dMv);
            (void)hipFree(p->dCost);
            p->dMv = nullptr;
            p->dCost = nullptr;
            p->cap = 0;
            if (!XG_HIP(hipMalloc((void **)&p->dMv, sizeof(int16_t) * 16 * (size_t)nb)) ||
                !XG_HIP(hipMalloc((void **)&p->dCost, sizeof(uint32_t) * 4 * (size_t)nb)))
                return XG_ERR_NOMEM;
            p->cap = nb;
        }
        p->nb = nb;
        p->fw = fw;
        p->fh = fh;
        session_view(s, &v);
        p->gridVersion = v.blocksVersion;
    }
    if ((rc = refine_prepare(s, &cfg->refine, pred)) != XG_OK) return rc;
    PipeArgs &A = p->args;
    A.pics = v.dPics;
    A.curSlot = 0;
    A.refSlot0 = 1;
    A.refSlot1 = 2;
    A.fw = fw;
    A.fh = fh;
    A.nbx = nbx;
    A.nb = nb;
    A.range = cfg->range;
    A.ctu = cfg->ctuSize;
    A.lambda = cfg->sadLambda;
    A.fullMv = p->dMv;
    A.predMv = p->dMv + 4 * (size_t)nb;
    A.intMv = p->dMv + 8 * (size_t)nb;
    A.qMv = p->dMv + 12 * (size_t)nb;
    A.intCost = p->dCost;
    A.qCost = p->dCost + 2 * (size_t)nb;
    A.blocks = v.dBlocks;
    int ev[2];
    const bool timed = session_timer_begin(s, 4, ev);
    if (mode == XG_MODE_BATCHED) {
        const bool ok = enqueue_pipe(p, v.stream, s, true);
        if (timed) session_timer_end(s, 4, ev);
        return ok ? XG_OK : XG_ERR_HIP;
    }
    uint64_t key = hmix(hmix((uint64_t)nb, (uint64_t)(uintptr_t)v.dPics), (uint64_t)(uintptr_t)v.dBlocks);
    key = hmix(key, ((uint64_t)cfg->range << 40) | ((uint64_t)cfg->ctuSize << 32) | cfg->sadLambda);
    const uint8_t *rb = reinterpret_cast<const uint8_t *>(&cfg->refine);
    for (size_t i = 0; i < sizeof(xg_refine_cfg); i++) key = hmix(key, rb[i]);
    key = hmix(key, (uint64_t)(uintptr_t)p->dMv);
    key = hmix(key, (uint64_t)(uintptr_t)p->dCost);
    // the captured refinement launch: grid, task lists, tile table, per-block state
    key = hmix(key, refine_prepared_key(s));
    if (!p->exec || key != p->key) {
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        p->exec = nullptr;
        p->graph = nullptr;
        if (!XG_HIP(hipStreamBeginCapture(v.stream, hipStreamCaptureModeThreadLocal))) return XG_ERR_HIP;
        const bool ok = enqueue_pipe(p, v.stream, s, false);
        hipGraph_t g = nullptr;
        const bool ended = XG_HIP(hipStreamEndCapture(v.stream, &g));
        if (!ok || !ended) return XG_ERR_HIP;
        p->graph = g;
        if (!XG_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0))) return XG_ERR_HIP;
        p->key = key;
    }
    const bool ok = XG_HIP(hipGraphLaunch(p->exec, v.stream));
    if (timed) session_timer_end(s, 4, ev);
    return ok ? XG_OK : XG_ERR_HIP;
}

int32_t xg_me_pipeline_read(xg_session *s, int16_t *fullMv, int16_t *predMv, int16_t *intMv, uint32_t *intCost,
                            int16_t *qMv, uint32_t *qCost)
{
    PipeState *p = s ? static_cast<PipeState *>(session_pipe(s)) : nullptr;
    if (!p || !p->nb) {
        set_error(XG_ERR_STATE, "xg_me_pipeline_read: no pipeline run on this session");
        return XG_ERR_STATE;
    }
    if (xg_session_sync(s) != XG_OK) return XG_ERR_HIP;
    const size_t n4 = 4 * (size_t)p->nb, n2 = 2 * (size_t)p->nb;
    bool ok = true;
    if (fullMv) ok &= XG_HIP(hipMemcpy(fullMv, p->args.fullMv, n4 * sizeof(int16_t), hipMemcpyDeviceToHost));
    if (predMv) ok &= XG_HIP(hipMemcpy(predMv, p->args.predMv, n4 * sizeof(int16_t), hipMemcpyDeviceToHost));
    if (intMv) ok &= XG_HIP(hipMemcpy(intMv, p->args.intMv, n4 * sizeof(int16_t), hipMemcpyDeviceToHost));
    if (qMv) ok &= XG_HIP(hipMemcpy(qMv, p->args.qMv, n4 * sizeof(int16_t), hipMemcpyDeviceToHost));
    if (intCost) ok &= XG_HIP(hipMemcpy(intCost, p->args.intCost, n2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (qCost) ok &= XG_HIP(hipMemcpy(qCost, p->args.qCost, n2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return ok ? XG_OK : XG_ERR_HIP;
}

}  // extern "C"

// xg_refine.hip -- batched fractional motion refinement fused with motion
// compensation.
//
// For every block and each of its ncand candidate MVs (base MV + offset,
// 1/16 pel, clipped to the CU MV window of
// h266_encoder/source/h266_enc_init.c:1462-1465) the luma 8-tap prediction
// (h266_core/source/h266_luma_mc.c, table dispatch of
// h266_encoder/source/h266_inter_prediction.c:143-172, filterIndex 0) and,
// when enabled, the Cb/Cr 4-tap predictions (h266_chroma_mc.c, address split
// of h266_inter_prediction.c:664-674) are formed and compared with the
// current block by SATD; cost = (SATD << 8) + (EG(dx) + EG(dy)) * lambda with
// the MVD in 1/16 pel.  The lowest cost wins (earliest candidate on ties) and
// its luma + chroma prediction is written into the prediction picture.
//
// Mapping (one wavefront = one workgroup of 64 lanes):
//  * a task is (block, first candidate, candidate count); tasks are grouped
//    per block shape and ordered largest first (per-shape work queues);
//  * the task's reference window -- every sample any of its candidates'
//    filters touch -- is staged once into LDS and reused by all candidates;
//  * lane = unit * (W / 2) + m: each lane walks columns 2m and 2m+1 of one
//    candidate (chroma: of one candidate's Cb or Cr) down the block, keeping
//    the first-pass values of the separable filter in registers as packed
//    16-bit row pairs, so the intermediate never leaves the VGPRs; both
//    filter passes are v_dot2_i32_i16 on sample pairs;
//  * SATD tiles (8x8, 4x4 or 2x2 as the block shape allows) are transformed
//    on packed int16: vertically in registers, horizontally inside the
//    register and across lanes with DPP; sums reduce within the wavefront;
//  * each task folds its best (cost << 32 | candidate) into a per-block
//    64-bit atomic minimum and counts its candidates into a per-block
//    counter; the task that completes a block performs the motion
//    compensation of the winner from a freshly staged LDS window and resets
//    the block's counters, so one launch does refinement and MC;
//  * batched mode: one persistent launch pulls tasks from an atomic queue
//    head, and the last workgroup to leave resets the queue, so a frame is a
//    single kernel node (graph-capturable, no memset); plain mode: one launch
//    per block shape.
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
constexpr int kPicSlots = kSessionPicSlots;
constexpr int kRegions = 8;        // spatial work queues, one per XCD of the MI355X

struct RefineArgs {
    const xg_picture *pics;
    const xg_block *blocks;
    const int2 *tasks;             // x: block index, y: c0 | (count << 16)
    int taskBegin, taskEnd;        // this launch's task range
    int *queue;                    // batched mode: [0..7] region queue heads, [8] exit count; nullptr in plain mode
    const int *regionBegin;        // batched mode: task offsets of the spatial queues (nregions + 1)
    int nregions;                  // batched mode: number of spatial queues (1..kRegions)
    uint32_t *candCost;
    xg_block_result *best;
    unsigned long long *bestKey;   // per block: min over tasks of (cost << 32 | cand); all-ones when idle
    int *done;                     // per block: candidates costed so far; 0 when idle
    int ncand;
    int offMinX, offMaxX, offMinY, offMaxY;
    uint32_t lambda;
    int ctuSize;
    int useChroma;
    const int16_t *off;            // device: offx[XG_MAX_CAND] then offy[XG_MAX_CAND]
    int doMc;                      // 0: costs only (tile timing), no block completion or MC
    const xg_picture *predTab;     // per picture slot: prediction planes (unpadded) the MC writes
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

// ---------------------------------------------------------------------------
// Interpolation + SATD of one plane for a pass of candidates, two adjacent
// columns per lane.
//
// Each lane owns columns 2m and 2m+1 of one unit (a candidate, or for chroma
// a candidate's Cb or Cr plane).  A window row is fetched as dword-aligned
// LDS words and realigned with v_alignbit, so every 16-bit sample pair
// (p[i], p[i+1]) the filters need is one register; the first pass is then
// TAPS/2 v_dot2_i32_i16 per column.  First-pass values are kept as packed
// consecutive-row pairs (t[r], t[r+1]), so the second pass is TAPS/2 more
// dot2.  Differences, the vertical Walsh-Hadamard stages and the first
// horizontal stages run on packed int16 (the largest magnitude before the
// last stage is 8 * 4 * 1023 < 32768); the last stage uses
// |a + b| + |a - b| = 2 * max(|a|, |b|), so no value ever needs 17 bits.
// ---------------------------------------------------------------------------
typedef short v2s __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) uint32_t gdword;

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
__device__ __forceinline__ uint32_t align16(uint32_t hi, uint32_t lo, uint32_t sh)
{
    return __builtin_amdgcn_alignbit(hi, lo, sh);
}

// Lane exchange of a packed pair with lane ^ M (M = 1, 2: DPP quad permutes).
template <int M>
__device__ __forceinline__ v2s xlane(v2s v)
{
    return as_v2s((uint32_t)shfl_xor<M>((int)as_u32(v)));
}

// Butterfly (a, b) -> (a + b, a - b) between the two halves of a register.
__device__ __forceinline__ v2s bfly_in(v2s v)
{
    const v2s sw = __builtin_shufflevector(v, v, 1, 0);
    return v * (v2s){1, -1} + sw;
}

// Butterfly across lanes: with a the value of the lane whose bit M is clear
// and b its partner's, that lane keeps a + b and the partner keeps a - b.
template <int M>
__device__ __forceinline__ v2s bfly_lane(v2s v, int lane)
{
    const short sg = (lane & M) ? -1 : 1;
    return v * (v2s){sg, sg} + xlane<M>(v);
}

__device__ __forceinline__ v2s absv(v2s v) { return __builtin_elementwise_max(v, (v2s){0, 0} - v); }

// The last horizontal stage (across lane ^ LM, or inside the register when
// LM == 0) is folded into the absolute sum: each lane adds max(|own|,
// |partner|) per half, so the pair together adds |a + b| + |a - b|.
// SATD (HM normalisation) of one TS x TS tile whose rows d[0..TS-1] hold, in
// each of the TS/2 lanes of the tile, the differences of two adjacent
// columns.  Every lane of the tile receives the tile's SATD.
template <int TS>
__device__ __forceinline__ int tile_satd2(v2s *d, int lane)
{
#pragma unroll
    for (int len = 1; len < TS; len <<= 1)          // vertical transform in registers
#pragma unroll
        for (int i = 0; i < TS; i += len << 1)
#pragma unroll
            for (int j = i; j < i + len; j++) {
                const v2s a = d[j], b = d[j + len];
                d[j] = a + b;
                d[j + len] = a - b;
            }
    // horizontal stages, each over all rows before the next (independent DPP
    // sources between a write and its cross-lane read)
    if constexpr (TS >= 4) {
#pragma unroll
        for (int r = 0; r < TS; r++) d[r] = bfly_in(d[r]);
    }
    if constexpr (TS >= 8) {
#pragma unroll
        for (int r = 0; r < TS; r++) d[r] = bfly_lane<1>(d[r], lane);
    }
    constexpr int LM = TS == 2 ? 0 : TS == 4 ? 1 : 2;
    int acc = 0;
#pragma unroll
    for (int r = 0; r < TS; r++) d[r] = absv(d[r]);
#pragma unroll
    for (int r = 0; r < TS; r++) {
        v2s mx;
        if constexpr (LM == 0) mx = __builtin_elementwise_max(d[r], __builtin_shufflevector(d[r], d[r], 1, 0));
        else mx = __builtin_elementwise_max(d[r], xlane<LM>(d[r]));
        acc = dot2(as_u32(mx), 0x00010001u, acc);
    }
    if constexpr (TS >= 4) acc += shfl_xor<1>(acc);
    if constexpr (TS >= 8) acc += shfl_xor<2>(acc);
    if constexpr (TS == 8) return (acc + 2) >> 2;
    else if constexpr (TS == 4) return (acc + 1) >> 1;
    else return acc;
}

// First-pass values of window row `row` (sample index of the lane's first
// column in the window, any parity) for the lane's two columns.
template <int TAPS>
__device__ __forceinline__ void hrow2(const uint16_t *win, int idx, const uint32_t *hp, int &t0, int &t1)
{
    const uint32_t *q = reinterpret_cast<const uint32_t *>(win) + (idx >> 1);
    const uint32_t sh = (uint32_t)(idx & 1) << 4;
    constexpr int ND = TAPS / 2 + 2;                // dwords covering TAPS + 1 samples at any parity
    uint32_t D[ND], E[ND - 1];
#pragma unroll
    for (int j = 0; j < ND; j++) D[j] = q[j];
#pragma unroll
    for (int j = 0; j < ND - 1; j++) E[j] = align16(D[j + 1], D[j], sh);      // (p[2j], p[2j+1])
    int s0 = kHOffset, s1 = kHOffset;
#pragma unroll
    for (int j = 0; j < TAPS / 2; j++) {
        s0 = dot2(E[j], hp[j], s0);
        s1 = dot2(align16(E[j + 1], E[j], 16), hp[j], s1);                  // (p[2j+1], p[2j+2])
    }
    t0 = s0 >> kHShift;
    t1 = s1 >> kHShift;
}

// One pass over up to 64 / (W / 2) units of one block.  unitPlane(u) and the
// window origin are supplied per lane by the caller.
//   win:   LDS window of the lane's plane; ls its row stride (even)
//   wx, wy: the lane's candidate's first tap position in the window
//   cur:   the plane's current block (global, dword aligned), cs its stride
template <int TAPS, int W, int TS>
__device__ __forceinline__ int interp_satd_pass(const uint16_t *win, int ls, int wx, int wy, int fx, int fy,
                                                const uint16_t *cur, intptr_t cs, int H, int m, int lane)
{
    uint32_t hp[TAPS / 2], vp[TAPS / 2];
    {
        const int8_t *ch = taps_for<TAPS>(fx, 0), *cv = taps_for<TAPS>(fy, 0);
#pragma unroll
        for (int j = 0; j < TAPS / 2; j++) {
            hp[j] = pack16(ch[2 * j], ch[2 * j + 1]);
            vp[j] = pack16(cv[2 * j], cv[2 * j + 1]);
        }
    }
    constexpr int NP = TS + TAPS - 2;               // packed row pairs per column in a tile
    uint32_t P0[NP], P1[NP];
    int tl0, tl1;
    const int base = wy * ls + wx + 2 * m;
    {
        int a0, a1;
        hrow2<TAPS>(win, base, hp, a0, a1);
#pragma unroll
        for (int r = 1; r < TAPS - 1; r++) {
            int b0, b1;
            hrow2<TAPS>(win, base + r * ls, hp, b0, b1);
            P0[r - 1] = pack16(a0, b0);
            P1[r - 1] = pack16(a1, b1);
            a0 = b0;
            a1 = b1;
        }
        tl0 = a0;
        tl1 = a1;
    }
    const gdword *gc = (const gdword *)(cur + 2 * m);
    const intptr_t csw = cs >> 1;
    int acc = 0;
    for (int y0 = 0; y0 < H; y0 += TS) {
        uint32_t cw[TS];
#pragma unroll
        for (int k = 0; k < TS; k++) cw[k] = gc[(intptr_t)(y0 + k) * csw];
#pragma unroll
        for (int k = 0; k < TS; k++) {
            int b0, b1;
            hrow2<TAPS>(win, base + (y0 + TAPS - 1 + k) * ls, hp, b0, b1);
            P0[TAPS - 2 + k] = pack16(tl0, b0);
            P1[TAPS - 2 + k] = pack16(tl1, b1);
            tl0 = b0;
            tl1 = b1;
        }
        v2s d[TS];
#pragma unroll
        for (int k = 0; k < TS; k++) {
            int s0 = kVOffset, s1 = kVOffset;
#pragma unroll
            for (int j = 0; j < TAPS / 2; j++) {
                s0 = dot2(P0[k + 2 * j], vp[j], s0);
                s1 = dot2(P1[k + 2 * j], vp[j], s1);
            }
            d[k] = as_v2s(cw[k]) - as_v2s(pack16(clip_pix(s0 >> kVShift), clip_pix(s1 >> kVShift)));
        }
        const int ts = tile_satd2<TS>(d, lane);
        if ((m & (TS / 2 - 1)) == 0) acc += ts;
#pragma unroll
        for (int k = 0; k < TAPS - 2; k++) {
            P0[k] = P0[TS + k];
            P1[k] = P1[TS + k];
        }
    }
    return acc;
}

// Luma SATD of every candidate of a task; lane = unit * (W / 2) + m.
template <int W, int TS>
__device__ __forceinline__ void luma_task(const RefineArgs &a, const xg_block &B, const MvWin &mw,
                                          const uint16_t *lds, int wx0, int wy0, const uint16_t *cur,
                                          intptr_t cs, int c0, int nc, uint32_t *csum, const int16_t *sOff)
{
    constexpr int L = W / 2, CP = 64 / L;
    const int lane = threadIdx.x, m = lane & (L - 1), u = lane / L;
    for (int base = 0; base < nc; base += CP) {
        const int ci = base + u;
        const int c = c0 + min(ci, nc - 1);
        const int mvx = clampi(B.mvx + sOff[c], mw.mnx, mw.mxx);
        const int mvy = clampi(B.mvy + sOff[XG_MAX_CAND + c], mw.mny, mw.mxy);
        const int rx = (B.x << 4) + mvx, ry = (B.y << 4) + mvy;
        int acc = interp_satd_pass<8, W, TS>(lds, kLumaLW, (rx >> 4) - 3 - wx0, (ry >> 4) - 3 - wy0, rx & 15,
                                             ry & 15, cur, cs, B.h, m, lane);
        acc = group_sum<L>(acc);
        if (m == 0 && ci < nc) csum[ci] = (uint32_t)acc;
    }
}

// Chroma SATD (Cb + Cr) of every candidate of a task; unit = 2 * candidate +
// plane, lane = unit * (WC / 2) + m, so a candidate's two planes are adjacent.
template <int WC, int TC>
__device__ __forceinline__ void chroma_task(const RefineArgs &a, const xg_block &B, const MvWin &mw,
                                            const uint16_t *ldsU, const uint16_t *ldsV, int cx0, int cy0,
                                            const uint16_t *curU, const uint16_t *curV, intptr_t cs, int c0,
                                            int nc, uint32_t *csum, const int16_t *sOff)
{
    constexpr int L = WC / 2 > 0 ? WC / 2 : 1, UP = 64 / L;
    const int lane = threadIdx.x, m = lane & (L - 1), u = lane / L;
    for (int base = 0; base < 2 * nc; base += UP) {
        const int ui = base + u;
        const int ci = ui >> 1, pl = ui & 1;
        const int c = c0 + min(ci, nc - 1);
        const int mvx = clampi(B.mvx + sOff[c], mw.mnx, mw.mxx);
        const int mvy = clampi(B.mvy + sOff[XG_MAX_CAND + c], mw.mny, mw.mxy);
        const int rx = (B.x << 4) + mvx, ry = (B.y << 4) + mvy;
        int acc = interp_satd_pass<4, WC, TC>(pl ? ldsV : ldsU, kChromaLW, (rx >> 5) - 1 - cx0,
                                              (ry >> 5) - 1 - cy0, rx & 31, ry & 31, pl ? curV : curU, cs,
                                              B.h >> 1, m, lane);
        acc = group_sum<2 * L>(acc);
        if (m == 0 && pl == 0 && ci < nc) csum[ci] += (uint32_t)acc;
    }
}

// Motion compensation of one plane from an LDS window whose first row and
// column are TAPS/2-1 samples above/left of the block's integer position.
// Lanes [lane0, lane0 + nlanes) split the block into column x row-chunk
// units; each unit walks its rows with a sliding window of first-pass values.
template <int TAPS>
__device__ __forceinline__ void mc_plane(const uint16_t *win, int ls, int fx, int fy, int W, int H,
                                         uint16_t *dst, intptr_t ds, int lane0, int nlanes)
{
    const int lid = (int)threadIdx.x - lane0;
    if (lid < 0 || lid >= nlanes) return;
    const int chunks = max(1, min(nlanes / W, H));
    const int x = lid % W, ck = lid / W;
    if (ck >= chunks) return;
    const int rp = H / chunks, y0 = ck * rp;
    int hc[TAPS], vc[TAPS];
    const int8_t *ch = taps_for<TAPS>(fx, 0), *cv = taps_for<TAPS>(fy, 0);
#pragma unroll
    for (int j = 0; j < TAPS; j++) {
        hc[j] = ch[j];
        vc[j] = cv[j];
    }
    const uint16_t *p = win + y0 * ls + x;
    int t[TAPS];
#pragma unroll
    for (int k = 0; k < TAPS - 1; k++) t[k] = hrow_lds<TAPS>(p + k * ls, hc);
    for (int y = 0; y < rp; y++) {
        t[TAPS - 1] = hrow_lds<TAPS>(p + (y + TAPS - 1) * ls, hc);
        int s = 0;
#pragma unroll
        for (int j = 0; j < TAPS; j++) s += vc[j] * t[j];
        dst[(intptr_t)(y0 + y) * ds + x] = (uint16_t)vfinal_pix(s);
#pragma unroll
        for (int k = 0; k < TAPS - 1; k++) t[k] = t[k + 1];
    }
}

// Best-candidate record and luma + chroma prediction of block b.
__device__ __forceinline__ void mc_block(const RefineArgs &a, int b, const xg_block &B, const MvWin &mw,
                                         unsigned long long key, uint16_t *lds, const int16_t *sOff)
{
    const int lane = threadIdx.x;
    const int c = (int)(uint32_t)key;
    const int mvx = clampi(B.mvx + sOff[c], mw.mnx, mw.mxx);
    const int mvy = clampi(B.mvy + sOff[XG_MAX_CAND + c], mw.mny, mw.mxy);
    if (lane == 0) {
        xg_block_result r;
        r.mvx = mvx;
        r.mvy = mvy;
        r.cost = (uint32_t)(key >> 32);
        r.cand = c;
        a.best[b] = r;
    }
    const xg_picture &R = a.pics[B.ref];
    const int rx = (B.x << 4) + mvx, ry = (B.y << 4) + mvy;
    {   // luma: (H+7) x (W+7) samples from 3 above / 3 left of the integer position
        const intptr_t rs = R.plane[0].stride;
        const uint16_t *src = R.plane[0].data + (intptr_t)((ry >> 4) - 3) * rs + ((rx >> 4) - 3);
        const int ww = B.w + 7, hh = B.h + 7;
        for (int r = 0; r < hh; r++)
            for (int cc = lane; cc < ww; cc += 64) lds[r * kLumaLW + cc] = src[(intptr_t)r * rs + cc];
        __syncthreads();
        const xg_plane &P0 = a.predTab[B.pic].plane[0];
        mc_plane<8>(lds, kLumaLW, rx & 15, ry & 15, B.w, B.h, P0.data + (intptr_t)B.y * P0.stride + B.x,
                    P0.stride, 0, 64);
        __syncthreads();
    }
    {   // chroma: (H/2+3) x (W/2+3) samples per plane from 1 above / 1 left
        const int wc = B.w >> 1, hc = B.h >> 1, ww = wc + 3, hh = hc + 3;
        uint16_t *lu = lds, *lv = lds + kChromaLW * kChromaLH;
        const intptr_t rs = R.plane[1].stride;
        const intptr_t so = (intptr_t)((ry >> 5) - 1) * rs + ((rx >> 5) - 1);
        for (int r = 0; r < hh; r++)
            for (int cc = lane; cc < 2 * ww; cc += 64) {
                const int pl = cc >= ww, col = cc - pl * ww;
                (pl ? lv : lu)[r * kChromaLW + col] = R.plane[1 + pl].data[so + (intptr_t)r * rs + col];
            }
        __syncthreads();
        const xg_plane &P1 = a.predTab[B.pic].plane[1], &P2 = a.predTab[B.pic].plane[2];
        const intptr_t o1 = (intptr_t)(B.y >> 1) * P1.stride + (B.x >> 1);
        const intptr_t o2 = (intptr_t)(B.y >> 1) * P2.stride + (B.x >> 1);
        mc_plane<4>(lu, kChromaLW, rx & 31, ry & 31, wc, hc, P1.data + o1, P1.stride, 0, 32);
        mc_plane<4>(lv, kChromaLW, rx & 31, ry & 31, wc, hc, P2.data + o2, P2.stride, 32, 32);
        __syncthreads();
    }
}

__device__ __forceinline__ void refine_one(const RefineArgs &a, int task, uint16_t *lds, uint32_t *csum,
                                           const int16_t *sOff, int *sLast, unsigned long long *sKey)
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

    // ---- costs, task-local best, block completion ----
    unsigned long long key = ~0ull;
    if (lane < nc) {
        const int c = c0 + lane;
        const int mvx = clampi(B.mvx + sOff[c], mw.mnx, mw.mxx);
        const int mvy = clampi(B.mvy + sOff[XG_MAX_CAND + c], mw.mny, mw.mxy);
        const uint32_t bits = eg_bits(mvx - B.mvpx) + eg_bits(mvy - B.mvpy);
        const uint32_t cost = (csum[lane] << 8) + bits * a.lambda;
        a.candCost[(size_t)b * a.ncand + c] = cost;
        key = ((unsigned long long)cost << 32) | (uint32_t)c;
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const unsigned long long o = (unsigned long long)__shfl_xor((long long)key, m, 64);
        key = o < key ? o : key;
    }
    if (lane == 0) {
        int last = 0;
        if (!a.doMc) {
            last = 0;
        } else if (nc == a.ncand) {
            last = 1;                                // the task covers the whole block
        } else {
            atomicMin(&a.bestKey[b], key);
            __threadfence();
            last = atomicAdd(&a.done[b], nc) + nc == a.ncand;
            if (last) {
                __threadfence();
                key = atomicExch(&a.bestKey[b], ~0ull);   // read and reset for the next launch
                atomicExch(&a.done[b], 0);
            }
        }
        *sLast = last;
        *sKey = key;
    }
    __syncthreads();
    if (*sLast) mc_block(a, b, B, mw, *sKey, lds, sOff);
    __syncthreads();
}

#define XG_REFINE_LDS                                        \
    __shared__ __align__(16) uint16_t lds[kLumaLW * kLumaLH]; \
    __shared__ uint32_t csum[64];                            \
    __shared__ int16_t sOff[2 * XG_MAX_CAND];                \
    __shared__ int s_task, s_last;                           \
    __shared__ unsigned long long s_key;                     \
    for (int i = threadIdx.x; i < 2 * XG_MAX_CAND; i += 64) sOff[i] = a.off[i]; \
    __syncthreads();

// Batched form: a persistent grid drains the per-region task queues.
__global__ void __launch_bounds__(64) k_refine_persistent(RefineArgs a)
{
    XG_REFINE_LDS
    // Workgroups are dispatched to the XCDs round-robin, so workgroup i runs on
    // XCD i % 8.  Each drains the queue of its own spatial region first --
    // neighbouring blocks share reference samples in that XCD's L2 -- then
    // steals from the other regions in turn.
    int region = (int)(blockIdx.x % kRegions) % a.nregions, exhausted = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            int task = -1;
            while (exhausted < a.nregions) {
                const int beg = a.regionBegin[region], cnt = a.regionBegin[region + 1] - beg;
                const int i = cnt > 0 ? atomicAdd(&a.queue[region], 1) : 0;
                if (i < cnt) {
                    task = beg + i;
                    break;
                }
                region = (region + 1) % a.nregions;
                exhausted++;
            }
            s_task = task;
        }
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_task);
        __syncthreads();
        if (t < 0) break;
        refine_one(a, t, lds, csum, sOff, &s_last, &s_key);
    }
    // the last workgroup to leave resets the queues for the next launch
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&a.queue[kRegions], 1) == (int)gridDim.x - 1) {
            for (int r = 0; r < a.nregions; r++) atomicExch(&a.queue[r], 0);
            atomicExch(&a.queue[kRegions], 0);
        }
    }
}

// Plain form: one launch per block shape, one workgroup per task.
__global__ void __launch_bounds__(64) k_refine_shape(RefineArgs a)
{
    XG_REFINE_LDS
    const int t = a.taskBegin + (int)blockIdx.x;
    if (t < a.taskEnd) refine_one(a, t, lds, csum, sOff, &s_last, &s_key);
}
#undef XG_REFINE_LDS

}  // namespace xg

// ---------------------------------------------------------------------------
// Session
// ---------------------------------------------------------------------------
using namespace xg;

struct xg_session {
    hipStream_t stream = nullptr;
    int maxBlocks = 0;
    int numCUs = 256;
    int persistFactor = 11;            // persistent workgroups per CU (LDS allows 11)
    int nregions = 1;                  // spatial work queues of the batched form

    xg_picture pics[kPicSlots];
    bool picSet[kPicSlots] = {};
    xg_picture preds[kPicSlots];       // prediction target of each current-picture slot
    bool predSet[kPicSlots] = {};
    xg_picture *dPred = nullptr;       // device copy of preds
    xg_picture predUploaded[kPicSlots];
    bool predValid = false;
    xg_picture *dPics = nullptr;

    std::vector<xg_block> blocks;
    xg_block *dBlocks = nullptr;
    int nblocks = 0;
    uint64_t blocksVersion = 1;
    int lastNcand = 0;                 // ncand of the last refinement

    // per-shape work queues
    int tile[kNumShapes];              // candidates per task
    std::vector<int2> tasks;
    int2 *dTasks = nullptr;            // shape-major (plain mode)
    int2 *dTasksR = nullptr;           // region-major, heaviest first within a region (batched mode)
    int *dRegionBegin = nullptr;
    size_t taskCap = 0;
    int shapeTaskBegin[kNumShapes + 1] = {};
    std::vector<int> shapeOrder;       // shapes, heaviest tasks first
    uint64_t tasksKey = 0;             // (blocksVersion, ncand, tile table) the tasks were built for

    uint32_t *dCandCost = nullptr;
    xg_block_result *dBest = nullptr;
    int16_t *dOff = nullptr;           // candidate offsets of the last configuration
    int16_t hOff[2 * XG_MAX_CAND];
    bool offValid = false;
    int *dQueue = nullptr;             // persistent work-queue head and exit count
    int *dDone = nullptr;              // per-block candidate counters
    unsigned long long *dBestKey = nullptr;  // per-block best (cost << 32 | cand)
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

bool session_view(xg_session *s, SessionView *v)
{
    if (!s || !v) return false;
    v->stream = s->stream;
    v->dPics = s->dPics;
    v->dBlocks = s->dBlocks;
    v->hBlocks = s->blocks.data();
    v->nblocks = s->nblocks;
    v->hPics = s->pics;
    v->picSet = s->picSet;
    v->numCUs = s->numCUs;
    return true;
}

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

// Region of a block for the XCD-local queues: horizontal bands of its picture.
static int region_of(const xg_session *s, const xg_block &B)
{
    const int h = std::max(1, s->pics[B.pic].plane[0].height);
    return std::min(s->nregions - 1, (int)((long)B.y * s->nregions / h));
}

static bool build_tasks(xg_session *s, int ncand)
{
    uint64_t key = mix(mix(s->blocksVersion, (uint64_t)ncand), 0);
    for (int i = 0; i < kNumShapes; i++) key = mix(key, (uint64_t)s->tile[i]);
    for (int i = 0; i < kPicSlots; i++) key = mix(key, s->picSet[i] ? (uint64_t)s->pics[i].plane[0].height : 0);
    key = mix(key, (uint64_t)s->nregions);
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
    std::vector<std::vector<int2>> perRegion(s->nregions);
    int st[kNumShapes + 1];
    for (int oi = 0; oi < kNumShapes; oi++) {
        const int sh = order[oi];
        st[oi] = (int)s->tasks.size();
        const int cpt = std::max(1, std::min(s->tile[sh], ncand));
        for (int b : per[sh]) {
            const int r = region_of(s, s->blocks[b]);
            for (int c0 = 0; c0 < ncand; c0 += cpt) {
                const int2 t = make_int2(b, c0 | (std::min(cpt, ncand - c0) << 16));
                s->tasks.push_back(t);
                perRegion[r].push_back(t);
            }
        }
    }
    st[kNumShapes] = (int)s->tasks.size();
    memcpy(s->shapeTaskBegin, st, sizeof(st));
    s->shapeOrder = order;
    std::vector<int2> tasksR;
    tasksR.reserve(s->tasks.size());
    int rb[kRegions + 1];
    for (int r = 0; r <= kRegions; r++) {
        rb[r] = (int)tasksR.size();
        if (r < s->nregions) tasksR.insert(tasksR.end(), perRegion[r].begin(), perRegion[r].end());
    }
    if (s->tasks.size() > s->taskCap) {
        if (!XG_HIP(hipStreamSynchronize(s->stream))) return false;
        if (s->dTasks) hipFree(s->dTasks);
        if (s->dTasksR) hipFree(s->dTasksR);
        s->dTasks = s->dTasksR = nullptr;
        s->taskCap = 0;
        size_t cap = std::max<size_t>(s->tasks.size(), 1024);
        if (!XG_HIP(hipMalloc((void **)&s->dTasks, cap * sizeof(int2))) ||
            !XG_HIP(hipMalloc((void **)&s->dTasksR, cap * sizeof(int2))))
            return false;
        s->taskCap = cap;
    }
    if (!s->tasks.empty() && !stage_upload(s, s->dTasks, s->tasks.data(), s->tasks.size() * sizeof(int2))) return false;
    if (!tasksR.empty() && !stage_upload(s, s->dTasksR, tasksR.data(), tasksR.size() * sizeof(int2))) return false;
    if (!stage_upload(s, s->dRegionBegin, rb, sizeof(rb))) return false;
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

static bool validate_blocks(xg_session *s, bool pred, int ctu)
{
    for (int i = 0; i < s->nblocks; i++) {
        const xg_block &B = s->blocks[i];
        if (B.pic >= kPicSlots || B.ref >= kPicSlots || !s->picSet[B.pic] || !s->picSet[B.ref]) {
            set_error(XG_ERR_ARG, "block %d: picture slot not set", i);
            return false;
        }
        const xg_picture &P = s->pics[B.pic], &R = s->pics[B.ref];
        if (B.x < 0 || B.y < 0 || B.x + B.w > P.plane[0].width || B.y + B.h > P.plane[0].height ||
            (B.x & 3) || (B.y & 3)) {
            set_error(XG_ERR_ARG, "block %d (%d,%d %dx%d) outside its picture or not on the 4-sample grid", i, B.x,
                      B.y, B.w, B.h);
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
            if (!s->predSet[B.pic]) {
                set_error(XG_ERR_ARG, "block %d: no prediction picture for slot %d", i, B.pic);
                return false;
            }
            const xg_picture &pp = s->preds[B.pic];
            for (int p = 0; p < 3; p++) {
                const int sub = p ? 1 : 0;
                const xg_plane &pl = pp.plane[p];
                if (!pl.data || pl.stride < pl.width || ((B.x + B.w) >> sub) > pl.width ||
                    ((B.y + B.h) >> sub) > pl.height) {
                    set_error(XG_ERR_ARG, "prediction plane %d does not cover block %d", p, i);
                    return false;
                }
            }
        }
    return true;
}

static bool fill_args(xg_session *s, const xg_refine_cfg *cfg, RefineArgs &ra, bool doMc)
{
    int16_t off[2 * XG_MAX_CAND] = {};
    memset(&ra, 0, sizeof(ra));
    ra.pics = s->dPics;
    ra.blocks = s->dBlocks;
    ra.tasks = s->dTasks;
    ra.candCost = s->dCandCost;
    ra.best = s->dBest;
    ra.bestKey = s->dBestKey;
    ra.done = s->dDone;
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
    ra.doMc = doMc;
    ra.predTab = s->dPred;
    if (doMc && (!s->predValid || memcmp(s->predUploaded, s->preds, sizeof(s->preds)) != 0)) {
        if (!stage_upload(s, s->dPred, s->preds, sizeof(s->preds))) return false;
        memcpy(s->predUploaded, s->preds, sizeof(s->preds));
        s->predValid = true;
    }
    if (!s->offValid || memcmp(off, s->hOff, sizeof(off)) != 0) {
        if (!stage_upload(s, s->dOff, off, sizeof(off))) return false;
        memcpy(s->hOff, off, sizeof(off));
        s->offValid = true;
    }
    return true;
}

// One persistent launch over every task (refinement + MC) on the session stream.
static bool enqueue_batched(xg_session *s, RefineArgs ra, bool timed)
{
    const int grid = std::min(s->numCUs * s->persistFactor, std::max(1, (int)s->tasks.size()));
    ra.queue = s->dQueue;
    ra.tasks = s->dTasksR;
    ra.regionBegin = s->dRegionBegin;
    ra.nregions = s->nregions;
    ra.taskBegin = 0;
    ra.taskEnd = (int)s->tasks.size();
    int e0 = -1, e1 = -1;
    if (timed) {
        e0 = ev_get(s);
        e1 = ev_get(s);
        if (e0 < 0 || e1 < 0) timed = false;
    }
    if (timed) hipEventRecord(s->evPool[e0], s->stream);
    k_refine_persistent<<<grid, 64, 0, s->stream>>>(ra);
    if (timed) {
        hipEventRecord(s->evPool[e1], s->stream);
        s->evPending[0].push_back({e0, e1});
    }
    return XG_HIP(hipGetLastError());
}

// One launch per block shape, one workgroup per task.
static bool enqueue_plain(xg_session *s, RefineArgs ra, bool timed)
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
        ra.queue = nullptr;
        ra.tasks = s->dTasks;
        ra.taskBegin = b0;
        ra.taskEnd = b1;
        k_refine_shape<<<b1 - b0, 64, 0, s->stream>>>(ra);
    }
    if (timed) {
        hipEventRecord(s->evPool[e1], s->stream);
        s->evPending[0].push_back({e0, e1});
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
              XG_HIP(hipMalloc((void **)&s->dPred, sizeof(xg_picture) * kPicSlots)) &&
              XG_HIP(hipMalloc((void **)&s->dBlocks, sizeof(xg_block) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dCandCost, sizeof(uint32_t) * (size_t)maxBlocks * XG_MAX_CAND)) &&
              XG_HIP(hipMalloc((void **)&s->dBest, sizeof(xg_block_result) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dQueue, 64 * sizeof(int))) &&
              XG_HIP(hipMalloc((void **)&s->dRegionBegin, (kRegions + 1) * sizeof(int))) &&
              XG_HIP(hipMemset(s->dRegionBegin, 0, (kRegions + 1) * sizeof(int))) &&
              XG_HIP(hipMalloc((void **)&s->dDone, sizeof(int) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dBestKey, sizeof(unsigned long long) * (size_t)maxBlocks)) &&
              XG_HIP(hipMemset(s->dQueue, 0, 64 * sizeof(int))) &&
              XG_HIP(hipMemset(s->dDone, 0, sizeof(int) * (size_t)maxBlocks)) &&
              XG_HIP(hipMemset(s->dBestKey, 0xFF, sizeof(unsigned long long) * (size_t)maxBlocks)) &&
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
    hipFree(s->dPred);
    hipFree(s->dBlocks);
    hipFree(s->dTasks);
    hipFree(s->dTasksR);
    hipFree(s->dRegionBegin);
    hipFree(s->dCandCost);
    hipFree(s->dBest);
    hipFree(s->dQueue);
    hipFree(s->dDone);
    hipFree(s->dBestKey);
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
            pl.stride < pl.width + 2 * pl.pad || pl.pad < 0 || (pl.stride & 1) || ((uintptr_t)pl.data & 3)) {
            set_error(XG_ERR_ARG, "xg_session_set_picture: plane %d malformed (strides even, samples 4-byte aligned)",
                      p);
            return XG_ERR_ARG;
        }
    }
    if (pic->plane[1].stride != pic->plane[2].stride || pic->plane[1].pad != pic->plane[2].pad) {
        set_error(XG_ERR_ARG, "xg_session_set_picture: Cb and Cr planes must share stride and padding");
        return XG_ERR_ARG;
    }
    s->pics[slot] = *pic;
    s->picSet[slot] = true;
    if (!stage_upload(s, s->dPics + slot, pic, sizeof(xg_picture))) return XG_ERR_HIP;
    return XG_OK;
}

int32_t xg_session_set_pred(xg_session *s, int32_t slot, const xg_picture *pred)
{
    if (!s || !pred || slot < 0 || slot >= kPicSlots) {
        set_error(XG_ERR_ARG, "xg_session_set_pred: bad arguments");
        return XG_ERR_ARG;
    }
    s->preds[slot] = *pred;
    s->predSet[slot] = true;
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
    if (n && (!XG_HIP(hipMemsetAsync(s->dDone, 0, sizeof(int) * (size_t)n, s->stream)) ||
              !XG_HIP(hipMemsetAsync(s->dBestKey, 0xFF, sizeof(unsigned long long) * (size_t)n, s->stream))))
        return XG_ERR_HIP;
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

int32_t xg_session_set_sched(xg_session *s, int32_t regions, int32_t wgPerCU)
{
    if (!s || regions < 1 || regions > kRegions || wgPerCU < 1 || wgPerCU > 64) {
        set_error(XG_ERR_ARG, "xg_session_set_sched: regions 1..%d, workgroups per CU 1..64", kRegions);
        return XG_ERR_ARG;
    }
    s->nregions = regions;
    s->persistFactor = wgPerCU;
    return XG_OK;
}

int32_t xg_session_get_tile(xg_session *s, int32_t lgW, int32_t lgH)
{
    if (!s || lgW < 2 || lgW > 6 || lgH < 2 || lgH > 6) return -1;
    return s->tile[(lgW - 2) * 5 + (lgH - 2)];
}

int32_t xg_refine_frac(xg_session *s, const xg_refine_cfg *cfg, const xg_picture *pred, int32_t mode)
{
    if (!s || !validate_cfg(cfg)) return XG_ERR_ARG;
    if (s->nblocks == 0) return XG_OK;
    if (pred)   // one prediction picture for every current-picture slot in use
        for (int i = 0; i < s->nblocks; i++) {
            s->preds[s->blocks[i].pic] = *pred;
            s->predSet[s->blocks[i].pic] = true;
        }
    if (!validate_blocks(s, true, cfg->ctuSize)) return XG_ERR_ARG;
    if (!build_tasks(s, cfg->ncand)) return XG_ERR_HIP;
    RefineArgs ra;
    if (!fill_args(s, cfg, ra, true)) return XG_ERR_HIP;
    s->lastNcand = cfg->ncand;
    if (mode == XG_MODE_PLAIN) return enqueue_plain(s, ra, s->timing) ? XG_OK : XG_ERR_HIP;
    if (mode == XG_MODE_BATCHED) return enqueue_batched(s, ra, s->timing) ? XG_OK : XG_ERR_HIP;
    if (mode != XG_MODE_GRAPH) {
        set_error(XG_ERR_ARG, "xg_refine_frac: unknown mode %d", mode);
        return XG_ERR_ARG;
    }
    // graph key: everything captured by value in the kernel arguments
    uint64_t key = s->tasksKey;
    key = mix(key, ((uint64_t)cfg->lambda << 32) | (uint32_t)cfg->ctuSize);
    key = mix(key, (uint64_t)cfg->useChroma);
    key = mix(key, ((uint64_t)s->nregions << 16) | (uint64_t)s->persistFactor);
    for (int i = 0; i < cfg->ncand; i++) key = mix(key, ((uint64_t)(uint16_t)cfg->offx[i] << 16) | (uint16_t)cfg->offy[i]);
    if (!s->graphExec || key != s->graphKey) {
        if (s->graphExec) hipGraphExecDestroy(s->graphExec);
        if (s->graph) hipGraphDestroy(s->graph);
        s->graphExec = nullptr;
        s->graph = nullptr;
        if (!XG_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal))) return XG_ERR_HIP;
        const bool ok = enqueue_batched(s, ra, false);
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

// Device time of `reps` batched launches (costs only, no MC), best of 3.
static float time_frame(xg_session *s, const xg_refine_cfg *cfg, int reps, hipEvent_t e0, hipEvent_t e1, bool *ok)
{
    *ok = false;
    if (!build_tasks(s, cfg->ncand)) return 0.f;
    RefineArgs ra;
    if (!fill_args(s, cfg, ra, false)) return 0.f;
    float best = 1e30f;
    for (int trial = 0; trial < 4; trial++) {
        if (!XG_HIP(hipEventRecord(e0, s->stream))) return 0.f;
        for (int r = 0; r < reps; r++)
            if (!enqueue_batched(s, ra, false)) return 0.f;
        if (!XG_HIP(hipEventRecord(e1, s->stream)) || !XG_HIP(hipEventSynchronize(e1))) return 0.f;
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (trial > 0) best = std::min(best, ms);   // the first trial warms up
    }
    *ok = true;
    return best;
}

int32_t xg_session_autotune(xg_session *s, const xg_refine_cfg *cfg, int32_t reps, const char *path)
{
    if (!s || !validate_cfg(cfg) || reps < 1) return XG_ERR_ARG;
    if (s->nblocks == 0) return XG_OK;
    if (!validate_blocks(s, false, cfg->ctuSize)) return XG_ERR_ARG;
    hipEvent_t e0, e1;
    if (!XG_HIP(hipEventCreate(&e0)) || !XG_HIP(hipEventCreate(&e1))) return XG_ERR_HIP;
    int32_t rc = XG_OK;
    bool ok = true;
    // 1. scheduling: spatial queues x persistent workgroups per CU, whole frame
    {
        float bestMs = 1e30f;
        int bestR = s->nregions, bestF = s->persistFactor;
        for (int r : {1, 2, 4, 8})
            for (int f : {8, 11, 16}) {
                s->nregions = r;
                s->persistFactor = f;
                const float ms = time_frame(s, cfg, reps, e0, e1, &ok);
                if (!ok) { rc = XG_ERR_HIP; break; }
                if (ms < bestMs) { bestMs = ms; bestR = r; bestF = f; }
            }
        s->nregions = bestR;
        s->persistFactor = bestF;
    }
    // 2. candidates per task: coordinate descent over the shapes present,
    //    heaviest total work first, each choice timed on the whole frame
    long work[kNumShapes] = {};
    for (int b = 0; b < s->nblocks && rc == XG_OK; b++) work[shape_of(s->blocks[b].w, s->blocks[b].h)] += (long)s->blocks[b].w * s->blocks[b].h;
    std::vector<int> order;
    for (int sh = 0; sh < kNumShapes; sh++)
        if (work[sh]) order.push_back(sh);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
    float cur = rc == XG_OK ? time_frame(s, cfg, reps, e0, e1, &ok) : 0.f;
    if (!ok) rc = XG_ERR_HIP;
    for (int sh : order) {
        if (rc != XG_OK) break;
        int bestT = s->tile[sh];
        for (int cpt = 1; cpt <= cfg->ncand && rc == XG_OK; cpt <<= 1) {
            if (cpt == bestT) continue;
            const int keep = s->tile[sh];
            s->tile[sh] = cpt;
            const float ms = time_frame(s, cfg, reps, e0, e1, &ok);
            if (!ok) { rc = XG_ERR_HIP; break; }
            if (ms < cur * 0.995f) { cur = ms; bestT = cpt; }
            s->tile[sh] = keep;
        }
        s->tile[sh] = bestT;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc == XG_OK && path) {
        FILE *f = fopen(path, "w");
        if (!f) {
            set_error(XG_ERR_ARG, "xg_session_autotune: cannot write %s", path);
            return XG_ERR_ARG;
        }
        fprintf(f, "# xg266 refine schedule: regions %d wg_per_cu %d\n", s->nregions, s->persistFactor);
        fprintf(f, "# lgW lgH candidates_per_task (ncand %d)\n", cfg->ncand);
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
        if (sscanf(line, "# xg266 refine schedule: regions %d wg_per_cu %d", &lw, &lh) == 2) {
            xg_session_set_sched(s, lw, lh);
            continue;
        }
        if (line[0] == '#') continue;
        if (sscanf(line, "%d %d %d", &lw, &lh, &t) == 3 && xg_session_set_tile(s, lw, lh, t) == XG_OK) n++;
    }
    fclose(f);
    return n > 0 ? XG_OK : XG_ERR_ARG;
}

}  // extern "C"

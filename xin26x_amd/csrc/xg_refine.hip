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
//  * batched mode: one persistent launch; the host assigns the tasks to its
//    workgroups ahead of time (heaviest first to the least loaded workgroup,
//    optionally in XCD-aligned picture bands), so scheduling costs no atomics
//    and a frame is a single kernel node (graph-capturable, no memset);
//    plain mode: one launch per block shape.
#include <algorithm>
#include <queue>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

// Blocks are refined in sub-blocks of at most kSub x kSub luma samples (a
// block with a 64-sample side is four or two of them, one after another;
// SATD tiles never cross a sub-block edge, so the sum is the same).  The LDS
// windows are sized for one sub-block, which keeps a workgroup at ~9 KB.
constexpr int kSub = 32;
constexpr int kLumaLW = kSub + 16;     // LDS row stride (samples) of the luma window: W + 8 + 8 at most
constexpr int kLumaLH = kSub + 16;     // rows: H + 7 + 8 at most, plus the one-row fetch-ahead
constexpr int kChromaLW = kSub / 2 + 8;  // per-plane chroma window stride: W/2 + 4 + 4 at most
constexpr int kChromaLH = kSub / 2 + 8;  // rows: H/2 + 3 + 4 at most, plus the fetch-ahead
constexpr int kMaxOff = 64;        // |candidate offset| limit, 1/16 pel
constexpr int kNumShapes = 25;     // (lgW-2)*5 + (lgH-2), W, H in 4..64
constexpr int kPicSlots = kSessionPicSlots;
constexpr int kRegions = 8;        // picture bands of the batched form: at most one per XCD of the MI355X

struct RefineArgs {
    const xg_picture *pics;
    const xg_block *blocks;
    const int2 *tasks;             // x: block index, y: c0 | (count << 16)
    int taskBegin, taskEnd;        // this launch's task range
    const int *wgBegin;            // batched mode: task range [wgBegin[g], wgBegin[g+1]) of workgroup g
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
    int2 *mvField;                 // when set: the winner's MV is also written to every 4x4 unit
    int mvStride;                  //   of the block in this field (units per row)
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

// Task, block, list and picture descriptors are the same for every lane, but
// the kernel also stores to global memory, so they are loaded into vector
// registers.  Read back through readfirstlane they live in scalar registers,
// and so does every address and loop bound derived from them.
template <class T>
__device__ __forceinline__ T uniform_load(const T *p)
{
    static_assert(sizeof(T) % 4 == 0, "descriptor of whole dwords");
    constexpr int N = sizeof(T) / 4;
    uint32_t w[N];
    __builtin_memcpy(w, p, sizeof(T));
#pragma unroll
    for (int k = 0; k < N; k++) w[k] = (uint32_t)__builtin_amdgcn_readfirstlane((int)w[k]);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---------------------------------------------------------------------------
// Interpolation + SATD of one plane for a pass of candidates, two adjacent
// columns per lane.
//
// Each lane owns columns 2m and 2m+1 of one unit (a candidate, or for chroma
// a candidate's Cb or Cr plane).  A window row is fetched as dword-aligned
// LDS words (TAPS/2 + 1 of them) and filtered with taps packed to match the
// parity of the lane's first sample (tap_row), so each column of the first
// pass is TAPS/2 + 1 v_dot2_i32_i16 with no realignment.  First-pass values are kept as packed
// consecutive-row pairs (t[r], t[r+1]), so the second pass is TAPS/2 more
// dot2.  Differences, the vertical Walsh-Hadamard stages and the first
// horizontal stages run on packed int16 (the largest magnitude before the
// last stage is 8 * 4 * 1023 < 32768); the last stage uses
// |a + b| + |a - b| = 2 * max(|a|, |b|), so no value ever needs 17 bits.
// ---------------------------------------------------------------------------
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

// Shared memory of one refinement workgroup.  The luma and both chroma
// windows have their own regions, so one staging round trip fills all three
// and the motion compensation of the winner reads them in place.
struct RefineLds {
    uint32_t luma[kLumaLH * kLumaLW / 2];
    uint32_t cb[kChromaLH * kChromaLW / 2];
    uint32_t cr[kChromaLH * kChromaLW / 2];
    uint32_t tapL[16 * 3 * 5];     // luma filter (index 0): tap rows A, B, A' per phase (tap_row)
    uint32_t tapC[32 * 3 * 3];     // chroma filter: the same per phase
    uint32_t csum[64];
    int16_t off[2 * XG_MAX_CAND];
    unsigned long long key;
    int last;
};

// Packed tap pairs of one phase of a TAPS-tap filter c, each tap times 64
// (pack_mid16, clip_hi16), as three rows of TP = TAPS / 2 + 1 dwords:
//   A  = (c0, c1) (c2, c3) ... (c[T-2], c[T-1]) (0, 0)
//   B  = (0, c0) (c1, c2) ... (c[T-1], 0)
//   A' = (0, 0) (c0, c1) ... (c[T-2], c[T-1])
// A window row held as dwords D[0..TP) from the even column at or left of
// the first tap gives a lane's two outputs (columns x, x + 1) with TP dot2
// each and no realignment: first tap even, s0 = D.A and s1 = D.B; first tap
// odd (D[0] holds the sample left of it), s0 = D.B and s1 = D.A'.  The
// vertical pass uses the first TAPS / 2 pairs of A.
template <int TAPS>
__device__ __forceinline__ uint32_t tap_row(const int8_t *c, int row, int j)
{
    constexpr int H = TAPS / 2;
    if (row == 0) return j < H ? pack16(64 * c[2 * j], 64 * c[2 * j + 1]) : 0u;
    if (row == 1) return pack16(j == 0 ? 0 : 64 * c[2 * j - 1], j == H ? 0 : 64 * c[2 * j]);
    return j == 0 ? 0u : pack16(64 * c[2 * j - 2], 64 * c[2 * j - 1]);
}

__device__ __forceinline__ void lds_init(RefineLds &L, const RefineArgs &a)
{
    const int lane = threadIdx.x;
    for (int i = lane; i < 16 * 15; i += 64) L.tapL[i] = tap_row<8>(c_luma[0][i / 15], (i % 15) / 5, i % 5);
    for (int i = lane; i < 32 * 9; i += 64) L.tapC[i] = tap_row<4>(c_chroma[i / 9], (i % 9) / 3, i % 3);
    for (int i = lane; i < 2 * XG_MAX_CAND; i += 64) L.off[i] = a.off[i];
    __syncthreads();
}

// The horizontal taps of a lane, chosen by the parity of its first tap
// (tap_row): X gives column x, Y column x + 1.
template <int TAPS>
struct HTaps {
    static constexpr int TP = TAPS / 2 + 1;
    uint32_t X[TP], Y[TP];
    __device__ __forceinline__ void load(const uint32_t *rows, bool odd)
    {
        const uint32_t *x = rows + (odd ? TP : 0), *y = rows + (odd ? 2 * TP : TP);
#pragma unroll
        for (int j = 0; j < TP; j++) {
            X[j] = x[j];
            Y[j] = y[j];
        }
    }
};

// First pass of one window row for the lane's two columns, split into the
// LDS fetch (TAPS / 2 + 1 dwords from the one holding the first tap) and the
// arithmetic, so a row can be fetched while the previous one is filtered.
template <int TAPS>
struct HRow {
    static constexpr int ND = TAPS / 2 + 1;
    uint32_t D[ND];
    __device__ __forceinline__ void fetch(const uint32_t *q)
    {
#pragma unroll
        for (int j = 0; j < ND; j++) D[j] = q[j];
    }
    // Scaled sums 64 * (sum + 32) of the two columns (offset-free first pass,
    // taps x 64): their bits 8..23 are the first-pass values (pack_mid16).
    __device__ __forceinline__ void sums(const HTaps<TAPS> &T, int &s0, int &s1) const
    {
        s0 = dot2_first(D[0], T.X[0], kHRound64);
        s1 = dot2_first(D[0], T.Y[0], kHRound64);
#pragma unroll
        for (int j = 1; j < ND; j++) {
            s0 = dot2(D[j], T.X[j], s0);
            s1 = dot2(D[j], T.Y[j], s1);
        }
    }
    __device__ __forceinline__ void filter(const HTaps<TAPS> &T, int &t0, int &t1) const
    {
        sums(T, t0, t1);
        t0 >>= 8;
        t1 >>= 8;
    }
};

template <int TAPS>
__device__ __forceinline__ void hrow2(const uint32_t *win, int idx, const HTaps<TAPS> &T, int &t0, int &t1)
{
    HRow<TAPS> h;
    h.fetch(win + (idx >> 1));
    h.filter(T, t0, t1);
}

// n rows of a lane's dword column, stride csw dwords, walked with one
// pointer: the empty asm hides the pointer's history, so the compiler does
// not keep a 64-bit offset per row in scalar registers (they spilled).
template <int N>
__device__ __forceinline__ void load_rows(const gdword *g, intptr_t csw, uint32_t *dst)
{
#pragma unroll
    for (int k = 0; k < N; k++) {
        dst[k] = *g;
        g += csw;
        asm volatile("" : "+v"(g));
    }
}

// One TS-row tile of interp_satd_pass: the first pass of its TS window rows
// into the new pairs N0/N1 (the pairs before them are the last TAPS - 2 of
// C0/C1), the second pass, and the SATD against the current rows cw.  When
// `pre` is set the next tile's current rows (gn) are loaded into cn meanwhile.
// Returns the tile's SATD on the lane that holds it (column pair 0 of the
// tile), 0 on the others.
template <int TAPS, int TS, int LS>
__device__ __forceinline__ int satd_tile(const uint32_t *&q, HRow<TAPS> &cur_row, HRow<TAPS> &next_row,
                                         const HTaps<TAPS> &ht, const uint32_t *vt, int &tl0, int &tl1,
                                         const uint32_t *C0, const uint32_t *C1, uint32_t *N0, uint32_t *N1,
                                         const uint32_t *cw, uint32_t *cn, bool pre, const gdword *gn,
                                         intptr_t csw, int m, int lane)
{
    constexpr int lw = LS >> 1, CR = TAPS - 2;
    if (pre) load_rows<TS>(gn, csw, cn);
#pragma unroll
    for (int k = 0; k < TS; k++) {
        int b0, b1;
        next_row.fetch(q + (k + 1) * lw);           // one row past the window at the very end: unused
        cur_row.sums(ht, b0, b1);
        cur_row = next_row;
        N0[k] = pack_mid16(tl0, b0);
        N1[k] = pack_mid16(tl1, b1);
        tl0 = b0;
        tl1 = b1;
    }
    q += TS * lw;
    // pair v of the tile: the carried ones, then the new ones; the second
    // pass's 2 TS dot2 chains are interleaved tap by tap (consecutive dot2 of
    // one chain would wait on each other)
    auto p0 = [&](int v) { return v < CR ? C0[TS - CR + v] : N0[v - CR]; };
    auto p1 = [&](int v) { return v < CR ? C1[TS - CR + v] : N1[v - CR]; };
    v2s d[TS];
    int s0[TS], s1[TS];
#pragma unroll
    for (int k = 0; k < TS; k++) {
        s0[k] = dot2_first(p0(k), vt[0], 0);
        s1[k] = dot2_first(p1(k), vt[0], 0);
    }
#pragma unroll
    for (int j = 1; j < TAPS / 2; j++)
#pragma unroll
        for (int k = 0; k < TS; k++) {
            s0[k] = dot2(p0(k + 2 * j), vt[j], s0[k]);
            s1[k] = dot2(p1(k + 2 * j), vt[j], s1[k]);
        }
#pragma unroll
    for (int k = 0; k < TS; k++) d[k] = as_v2s(cw[k]) - as_v2s(pack_hi16(clip_hi16(s0[k]), clip_hi16(s1[k])));
    const int ts = tile_satd2<TS>(d, lane);
    return (m & (TS / 2 - 1)) == 0 ? ts : 0;
}

// Interpolation + SATD of one unit (a candidate's plane) of W x H samples for
// the lane's columns 2m, 2m+1:
//   win, LS: the plane's LDS window and its row stride in samples (even)
//   wx, wy:  window position of the unit's first tap
//   hp, vp:  tap rows of the horizontal and vertical phases (tap_row, LDS)
//   cur, cs: the current block (global, dword aligned) and its stride
// Window rows are fetched one row ahead of the row being filtered, current
// rows one tile ahead.
template <int TAPS, int W, int TS, int LS>
__device__ __forceinline__ int interp_satd_pass(const uint32_t *win, int wx, int wy, const uint32_t *hp,
                                                const uint32_t *vp, const uint16_t *cur, intptr_t cs, int H, int m,
                                                int lane)
{
    constexpr int NP = TS + TAPS - 2;               // packed row pairs per column in a tile
    constexpr int ls = LS, lw = LS >> 1;            // window row stride in samples and dwords (immediate offsets)
    const int base = wy * ls + wx + 2 * m;
    const uint32_t *q = win + (base >> 1);          // dword of the first tap of the current row
    HTaps<TAPS> ht;
    ht.load(hp, base & 1);
    uint32_t vt[TAPS / 2];                          // vertical taps x 64 (clip_hi16)
#pragma unroll
    for (int j = 0; j < TAPS / 2; j++) vt[j] = vp[j];
    uint32_t P0[NP], P1[NP];
    int tl0, tl1;
    HRow<TAPS> cur_row, next_row;
    cur_row.fetch(q);
    {
        int a0, a1;                                 // scaled first-pass sums (HRow::sums)
        next_row.fetch(q + lw);
        cur_row.sums(ht, a0, a1);
        cur_row = next_row;
#pragma unroll
        for (int r = 1; r < TAPS - 1; r++) {
            int b0, b1;
            next_row.fetch(q + (r + 1) * lw);
            cur_row.sums(ht, b0, b1);
            cur_row = next_row;
            P0[r - 1] = pack_mid16(a0, b0);
            P1[r - 1] = pack_mid16(a1, b1);
            a0 = b0;
            a1 = b1;
        }
        tl0 = a0;
        tl1 = a1;
    }
    q += (TAPS - 1) * lw;                           // cur_row holds row TAPS - 1
    const gdword *gc = (const gdword *)(cur + 2 * m);
    const intptr_t csw = cs >> 1;
    if constexpr (TS >= TAPS - 2) {
        // Two tiles per iteration with the row pairs and current rows in
        // ping-pong arrays: a tile writes its TS new pairs into one array and
        // the next tile takes the TAPS - 2 pairs it carries from that
        // array's end, so no registers are copied between tiles.
        constexpr int CR = TAPS - 2;
        uint32_t A0[TS], A1[TS], B0[TS], B1[TS], cwA[TS], cwB[TS];
#pragma unroll
        for (int k = 0; k < CR; k++) {
            B0[TS - CR + k] = P0[k];
            B1[TS - CR + k] = P1[k];
        }
        load_rows<TS>(gc, csw, cwA);
        int acc = 0;
        for (int y0 = 0;;) {
            acc += satd_tile<TAPS, TS, LS>(q, cur_row, next_row, ht, vt, tl0, tl1, B0, B1, A0, A1, cwA, cwB,
                                           y0 + TS < H, gc + (intptr_t)(y0 + TS) * csw, csw, m, lane);
            y0 += TS;
            if (y0 >= H) break;
            acc += satd_tile<TAPS, TS, LS>(q, cur_row, next_row, ht, vt, tl0, tl1, A0, A1, B0, B1, cwB, cwA,
                                           y0 + TS < H, gc + (intptr_t)(y0 + TS) * csw, csw, m, lane);
            y0 += TS;
            if (y0 >= H) break;
        }
        return acc;
    } else {
        uint32_t cw[TS], cn[TS];
        load_rows<TS>(gc, csw, cw);
        int acc = 0;
        for (int y0 = 0; y0 < H; y0 += TS) {
            if (y0 + TS < H)                            // the next tile's current rows, in flight meanwhile
                load_rows<TS>(gc + (intptr_t)(y0 + TS) * csw, csw, cn);
#pragma unroll
            for (int k = 0; k < TS; k++) {
                int b0, b1;
                next_row.fetch(q + (k + 1) * lw);       // one row past the window at the very end: unused
                cur_row.sums(ht, b0, b1);
                cur_row = next_row;
                P0[TAPS - 2 + k] = pack_mid16(tl0, b0);
                P1[TAPS - 2 + k] = pack_mid16(tl1, b1);
                tl0 = b0;
                tl1 = b1;
            }
            q += TS * lw;
            // second pass of the tile's rows, the 2 TS dot2 chains interleaved tap by tap
            // (consecutive dot2 of one chain would wait on each other)
            v2s d[TS];
            int s0[TS], s1[TS];
#pragma unroll
            for (int k = 0; k < TS; k++) {
                s0[k] = dot2_first(P0[k], vt[0], 0);
                s1[k] = dot2_first(P1[k], vt[0], 0);
            }
#pragma unroll
            for (int j = 1; j < TAPS / 2; j++)
#pragma unroll
                for (int k = 0; k < TS; k++) {
                    s0[k] = dot2(P0[k + 2 * j], vt[j], s0[k]);
                    s1[k] = dot2(P1[k + 2 * j], vt[j], s1[k]);
                }
#pragma unroll
            for (int k = 0; k < TS; k++)
                d[k] = as_v2s(cw[k]) - as_v2s(pack_hi16(clip_hi16(s0[k]), clip_hi16(s1[k])));
            const int ts = tile_satd2<TS>(d, lane);
            if ((m & (TS / 2 - 1)) == 0) acc += ts;
#pragma unroll
            for (int k = 0; k < TAPS - 2; k++) {
                P0[k] = P0[TS + k];
                P1[k] = P1[TS + k];
            }
#pragma unroll
            for (int k = 0; k < TS; k++) cw[k] = cn[k];
        }
        return acc;
    }
}

// Luma SATD of every candidate of a task; lane = unit * (W / 2) + m.
// (ox, oy): the sub-block's offset inside the block, H its height; csum accumulates.
template <int W, int TS>
__device__ __forceinline__ void luma_task(const xg_block &B, const MvWin &mw, const RefineLds &L, int wx0, int wy0,
                                          const uint16_t *cur, intptr_t cs, int H, int ox, int oy, int c0, int nc,
                                          uint32_t *csum)
{
    constexpr int LN = W / 2, CP = 64 / LN;
    const int lane = threadIdx.x, m = lane & (LN - 1), u = lane / LN;
    for (int base = 0; base < nc; base += CP) {
        const int ci = base + u;
        const int c = c0 + min(ci, nc - 1);
        const int mvx = clampi(B.mvx + L.off[c], mw.mnx, mw.mxx);
        const int mvy = clampi(B.mvy + L.off[XG_MAX_CAND + c], mw.mny, mw.mxy);
        const int rx = ((B.x + ox) << 4) + mvx, ry = ((B.y + oy) << 4) + mvy;
        int acc = interp_satd_pass<8, W, TS, kLumaLW>(L.luma, (rx >> 4) - 3 - wx0, (ry >> 4) - 3 - wy0,
                                             L.tapL + 15 * (rx & 15), L.tapL + 15 * (ry & 15), cur, cs, H, m, lane);
        acc = group_sum<LN>(acc);
        if (m == 0 && ci < nc) csum[ci] += (uint32_t)acc;
    }
}

// Chroma SATD (Cb + Cr) of every candidate of a task; unit = 2 * candidate +
// plane, lane = unit * (WC / 2) + m, so a candidate's two planes are adjacent.
template <int WC, int TC>
__device__ __forceinline__ void chroma_task(const xg_block &B, const MvWin &mw, const RefineLds &L, int cx0, int cy0,
                                            const uint16_t *curU, const uint16_t *curV, intptr_t cs, int HC, int ox,
                                            int oy, int c0, int nc, uint32_t *csum)
{
    constexpr int LN = WC / 2 > 0 ? WC / 2 : 1, UP = 64 / LN;
    const int lane = threadIdx.x, m = lane & (LN - 1), u = lane / LN;
    for (int base = 0; base < 2 * nc; base += UP) {
        const int ui = base + u;
        const int ci = ui >> 1, pl = ui & 1;
        const int c = c0 + min(ci, nc - 1);
        const int mvx = clampi(B.mvx + L.off[c], mw.mnx, mw.mxx);
        const int mvy = clampi(B.mvy + L.off[XG_MAX_CAND + c], mw.mny, mw.mxy);
        const int rx = ((B.x + ox) << 4) + mvx, ry = ((B.y + oy) << 4) + mvy;
        int acc = interp_satd_pass<4, WC, TC, kChromaLW>(pl ? L.cr : L.cb, (rx >> 5) - 1 - cx0, (ry >> 5) - 1 - cy0,
                                              L.tapC + 9 * (rx & 31), L.tapC + 9 * (ry & 31), pl ? curV : curU, cs,
                                              HC, m, lane);
        acc = group_sum<2 * LN>(acc);
        if (m == 0 && pl == 0 && ci < nc) csum[ci] += (uint32_t)acc;
    }
}

// Dword pointer to sample (x, y) of a plane; x even.
__device__ __forceinline__ const gdword *plane_dw(const xg_plane &pl, int x, int y)
{
    return (const gdword *)(pl.data + (intptr_t)y * pl.stride + x);
}

// Global -> LDS copy plan of one plane's window: `rows` rows of `wd` dwords
// (wd <= 64) from s (row pitch sp dwords) to d (row pitch dp dwords).  Lane
// (c, r0) = (lane % wd, lane / wd) copies column c of rows r0, r0 + per, ...
// (per = 64 / wd rows at a time).  Loads go in batches of 8 per lane, all
// issued before the batch's first store, and a lane loads only its own rows.
struct StagePlan {
    const gdword *s;     // the lane's first source dword
    uint32_t *d;         // its first LDS dword
    intptr_t sstep;      // source dwords from one of its rows to the next
    int dstep;           // LDS dwords likewise
    int n;               // its rows
    int batches;         // batches of the lanes with the most rows (wave-uniform)
};

__device__ __forceinline__ StagePlan stage_plan(uint32_t *d, int dp, const gdword *s, intptr_t sp, int rows, int wd)
{
    const int per = 64 / wd, c = (int)threadIdx.x % wd, r0 = (int)threadIdx.x / wd;
    StagePlan P;
    P.s = s + (intptr_t)r0 * sp + c;
    P.d = d + r0 * dp + c;
    P.sstep = (intptr_t)per * sp;
    P.dstep = per * dp;
    P.n = r0 < per && r0 < rows ? (rows - r0 + per - 1) / per : 0;
    P.batches = ((rows + per - 1) / per + 7) / 8;
    return P;
}

__device__ __forceinline__ void stage_load(const StagePlan &P, int k, uint32_t *v)
{
    const gdword *g = P.s + (intptr_t)(8 * k) * P.sstep;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        if (8 * k + j < P.n) v[j] = *g;
        g += P.sstep;
        asm volatile("" : "+v"(g));              // one 64-bit add per row, no offsets held per row
    }
}

__device__ __forceinline__ void stage_store(const StagePlan &P, int k, const uint32_t *v)
{
#pragma unroll
    for (int j = 0; j < 8; j++)
        if (8 * k + j < P.n) P.d[(8 * k + j) * P.dstep] = v[j];
}

// Motion compensation of one plane from its LDS window: lanes [lane0,
// lane0 + nlanes) split the W x H block into row groups of W / 2 lanes; each
// lane forms columns 2m, 2m+1 of its rows and stores them as one dword.
template <int TAPS>
__device__ __forceinline__ void mc_pass(const uint32_t *win, int ls, int wx, int wy, const uint32_t *hp, int fy, int W,
                                        int H, uint16_t *dst, intptr_t ds, int lane0, int nlanes)
{
    const int lid = (int)threadIdx.x - lane0;
    const int LN = W > 2 ? W >> 1 : 1;
    const int groups = min(nlanes / LN, H);
    const int m = lid & (LN - 1), g = lid / LN;
    if (lid < 0 || lid >= nlanes || g >= groups) return;
    const int rows = H / groups, y0 = g * rows;
    int vc[TAPS];
    const int8_t *cv = taps_for<TAPS>(fy, 0);
#pragma unroll
    for (int j = 0; j < TAPS; j++) vc[j] = cv[j];
    int t0[TAPS], t1[TAPS];
    const int base = (wy + y0) * ls + wx + 2 * m;
    HTaps<TAPS> ht;
    ht.load(hp, base & 1);
#pragma unroll
    for (int k = 0; k < TAPS - 1; k++) hrow2<TAPS>(win, base + k * ls, ht, t0[k], t1[k]);
    uint16_t *o = dst + (intptr_t)y0 * ds + 2 * m;
    for (int y = 0; y < rows; y++) {
        hrow2<TAPS>(win, base + (y + TAPS - 1) * ls, ht, t0[TAPS - 1], t1[TAPS - 1]);
        int s0 = 0, s1 = 0;                         // offset-free second pass (kHRound)
#pragma unroll
        for (int j = 0; j < TAPS; j++) {
            s0 += vc[j] * t0[j];
            s1 += vc[j] * t1[j];
        }
        *reinterpret_cast<uint32_t *>(o + (intptr_t)y * ds) = pack16(clip_pix(s0 >> kVShift), clip_pix(s1 >> kVShift));
#pragma unroll
        for (int k = 0; k < TAPS - 1; k++) {
            t0[k] = t0[k + 1];
            t1[k] = t1[k + 1];
        }
    }
}

// Windows of one sub-block (luma origin sx, sy, size sw x sh) for every
// candidate of the configuration: every luma / chroma sample any candidate's
// filters touch, from the even column at or left of the first tap
// (dword-aligned rows), staged into LDS.
struct SubWin {
    int wx0, wy0, cx0, cy0;
};

__device__ __forceinline__ SubWin stage_sub(const RefineArgs &a, const xg_block &B, const MvWin &mw,
                                            const xg_picture &R, int sx, int sy, int sw, int sh, RefineLds &L)
{
    const int qx0 = (sx << 4) + clampi(B.mvx + a.offMinX, mw.mnx, mw.mxx);
    const int qx1 = (sx << 4) + clampi(B.mvx + a.offMaxX, mw.mnx, mw.mxx);
    const int qy0 = (sy << 4) + clampi(B.mvy + a.offMinY, mw.mny, mw.mxy);
    const int qy1 = (sy << 4) + clampi(B.mvy + a.offMaxY, mw.mny, mw.mxy);
    SubWin w;
    w.wx0 = ((qx0 >> 4) - 3) & ~1;
    w.wy0 = (qy0 >> 4) - 3;
    const int WW = (qx1 >> 4) + sw + 4 - w.wx0, HH = (qy1 >> 4) - (qy0 >> 4) + sh + 7;
    w.cx0 = ((qx0 >> 5) - 1) & ~1;
    w.cy0 = (qy0 >> 5) - 1;
    const int CWW = (qx1 >> 5) + (sw >> 1) + 2 - w.cx0, CHH = (qy1 >> 5) - (qy0 >> 5) + (sh >> 1) + 3;
    const StagePlan pl = stage_plan(L.luma, kLumaLW / 2, plane_dw(R.plane[0], w.wx0, w.wy0), R.plane[0].stride >> 1,
                                    HH, (WW + 1) >> 1);
    const StagePlan pu = stage_plan(L.cb, kChromaLW / 2, plane_dw(R.plane[1], w.cx0, w.cy0), R.plane[1].stride >> 1,
                                    CHH, (CWW + 1) >> 1);
    const StagePlan pv = stage_plan(L.cr, kChromaLW / 2, plane_dw(R.plane[2], w.cx0, w.cy0), R.plane[2].stride >> 1,
                                    CHH, (CWW + 1) >> 1);
    const int nb = max(pl.batches, pu.batches);
    for (int k = 0; k < nb; k++) {
        uint32_t vl[8], vu[8], vv[8];
        if (k < pl.batches) stage_load(pl, k, vl);
        if (k < pu.batches) {
            stage_load(pu, k, vu);
            stage_load(pv, k, vv);
        }
        if (k < pl.batches) stage_store(pl, k, vl);
        if (k < pu.batches) {
            stage_store(pu, k, vu);
            stage_store(pv, k, vv);
        }
    }
    return w;
}

// Luma + chroma prediction of one sub-block of the winner from the staged windows.
__device__ __forceinline__ void mc_sub(const RefineArgs &a, const xg_block &B, int mvx, int mvy, int sx, int sy,
                                       int sw, int sh, const SubWin &w, const RefineLds &L)
{
    const xg_picture PT = uniform_load(&a.predTab[B.pic]);
    const xg_plane &P0 = PT.plane[0], &P1 = PT.plane[1], &P2 = PT.plane[2];
    const int rx = (sx << 4) + mvx, ry = (sy << 4) + mvy;   // the sub-block's reference position
    mc_pass<8>(L.luma, kLumaLW, (rx >> 4) - 3 - w.wx0, (ry >> 4) - 3 - w.wy0, L.tapL + 15 * (rx & 15), ry & 15, sw, sh,
               P0.data + (intptr_t)sy * P0.stride + sx, P0.stride, 0, 64);
    const int wc = sw >> 1, hc = sh >> 1, ox = (rx >> 5) - 1 - w.cx0, oy = (ry >> 5) - 1 - w.cy0;
    const uint32_t *hc4 = L.tapC + 9 * (rx & 31);
    mc_pass<4>(L.cb, kChromaLW, ox, oy, hc4, ry & 31, wc, hc,
               P1.data + (intptr_t)(sy >> 1) * P1.stride + (sx >> 1), P1.stride, 0, 64);
    mc_pass<4>(L.cr, kChromaLW, ox, oy, hc4, ry & 31, wc, hc,
               P2.data + (intptr_t)(sy >> 1) * P2.stride + (sx >> 1), P2.stride, 0, 64);
}

// Best-candidate record and luma + chroma prediction of block b.  A block of
// one sub-block predicts from the windows its task staged; a larger block
// stages each sub-block again (they hold every candidate's footprint).
__device__ __forceinline__ void mc_block(const RefineArgs &a, int b, const xg_block &B, const MvWin &mw,
                                         unsigned long long key, RefineLds &L, const SubWin &last)
{
    const int lane = threadIdx.x;
    const int c = uniform((int)(uint32_t)key);
    const int mvx = clampi(B.mvx + L.off[c], mw.mnx, mw.mxx);
    const int mvy = clampi(B.mvy + L.off[XG_MAX_CAND + c], mw.mny, mw.mxy);
    if (lane == 0) {
        xg_block_result r;
        r.mvx = mvx;
        r.mvy = mvy;
        r.cost = (uint32_t)(key >> 32);
        r.cand = c;
        a.best[b] = r;
    }
    if (a.mvField) {
        const int uw = B.w >> 2, nu = uw * (B.h >> 2);
        int2 *f = a.mvField + (intptr_t)(B.y >> 2) * a.mvStride + (B.x >> 2);
        for (int u = lane; u < nu; u += 64) f[(intptr_t)(u / uw) * a.mvStride + u % uw] = make_int2(mvx, mvy);
    }
    const int sw = min(B.w, kSub), sh = min(B.h, kSub);
    if (sw == B.w && sh == B.h) {
        mc_sub(a, B, mvx, mvy, B.x, B.y, sw, sh, last, L);
        return;
    }
    const xg_picture R = uniform_load(&a.pics[B.ref]);
    for (int sy = B.y; sy < B.y + B.h; sy += sh)
        for (int sx = B.x; sx < B.x + B.w; sx += sw) {
            __syncthreads();
            const SubWin w = stage_sub(a, B, mw, R, sx, sy, sw, sh, L);
            __syncthreads();
            mc_sub(a, B, mvx, mvy, sx, sy, sw, sh, w, L);
        }
}

__device__ __forceinline__ void refine_one(const RefineArgs &a, int2 tk, const xg_block &B, RefineLds &L)
{
    const int lane = threadIdx.x;
    const int b = tk.x, c0 = tk.y & 0xFFFF, nc = tk.y >> 16;
    const xg_picture P = uniform_load(&a.pics[B.pic]);
    const xg_picture R = uniform_load(&a.pics[B.ref]);
    const MvWin mw = mv_window(B, P.plane[0].width, P.plane[0].height, a.ctuSize);

    const intptr_t cs = P.plane[0].stride, ccs = P.plane[1].stride;
    L.csum[lane] = 0;
    uint32_t *csum = L.csum;
    const int sw = min(B.w, kSub), sh = min(B.h, kSub);
    SubWin w{};
    for (int sy = B.y; sy < B.y + B.h; sy += sh)
        for (int sx = B.x; sx < B.x + B.w; sx += sw) {
            __syncthreads();                           // the previous sub-block's readers are done
            w = stage_sub(a, B, mw, R, sx, sy, sw, sh, L);
            __syncthreads();
            const uint16_t *cur = P.plane[0].data + (intptr_t)sy * cs + sx;
            const int ox = sx - B.x, oy = sy - B.y;     // sub-block offset inside the block
            const bool t4 = sw == 4 || sh == 4;
            switch (sw) {
            case 4: luma_task<4, 4>(B, mw, L, w.wx0, w.wy0, cur, cs, sh, ox, oy, c0, nc, csum); break;
            case 8: t4 ? luma_task<8, 4>(B, mw, L, w.wx0, w.wy0, cur, cs, sh, ox, oy, c0, nc, csum)
                       : luma_task<8, 8>(B, mw, L, w.wx0, w.wy0, cur, cs, sh, ox, oy, c0, nc, csum); break;
            case 16: t4 ? luma_task<16, 4>(B, mw, L, w.wx0, w.wy0, cur, cs, sh, ox, oy, c0, nc, csum)
                        : luma_task<16, 8>(B, mw, L, w.wx0, w.wy0, cur, cs, sh, ox, oy, c0, nc, csum); break;
            default: t4 ? luma_task<32, 4>(B, mw, L, w.wx0, w.wy0, cur, cs, sh, ox, oy, c0, nc, csum)
                        : luma_task<32, 8>(B, mw, L, w.wx0, w.wy0, cur, cs, sh, ox, oy, c0, nc, csum); break;
            }
            __syncthreads();

            if (a.useChroma) {
                const intptr_t co = (intptr_t)(sy >> 1) * ccs + (sx >> 1);
                const uint16_t *cu = P.plane[1].data + co, *cv = P.plane[2].data + co;
                const int wc = sw >> 1, hc = sh >> 1;
                const int tc = (wc % 8 == 0 && hc % 8 == 0) ? 8 : (wc % 4 == 0 && hc % 4 == 0) ? 4 : 2;
                const int cx0 = w.cx0, cy0 = w.cy0;
                switch (wc) {
                case 2: chroma_task<2, 2>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum); break;
                case 4: tc == 4 ? chroma_task<4, 4>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum)
                                : chroma_task<4, 2>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum); break;
                case 8: tc == 8 ? chroma_task<8, 8>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum)
                      : tc == 4 ? chroma_task<8, 4>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum)
                                : chroma_task<8, 2>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum); break;
                default: tc == 8 ? chroma_task<16, 8>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum)
                       : tc == 4 ? chroma_task<16, 4>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum)
                                 : chroma_task<16, 2>(B, mw, L, cx0, cy0, cu, cv, ccs, hc, ox, oy, c0, nc, csum); break;
                }
                __syncthreads();
            }
        }

    // ---- costs, task-local best, block completion ----
    unsigned long long key = ~0ull;
    if (lane < nc) {
        const int c = c0 + lane;
        const int mvx = clampi(B.mvx + L.off[c], mw.mnx, mw.mxx);
        const int mvy = clampi(B.mvy + L.off[XG_MAX_CAND + c], mw.mny, mw.mxy);
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
        L.last = last;
        L.key = key;
    }
    __syncthreads();
    if (uniform(L.last)) mc_block(a, b, B, mw, L.key, L, w);
    __syncthreads();
}

// Batched form: a persistent grid; workgroup g runs the tasks the host
// assigned to it (balanced by estimated cost, heaviest first), so the launch
// takes no atomics to schedule and leaves no state behind.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_refine_persistent(RefineArgs a)
{
    __shared__ RefineLds L;
    lds_init(L, a);
    const int e = uniform(a.wgBegin[blockIdx.x + 1]);
    for (int t = uniform(a.wgBegin[blockIdx.x]); t < e; t++) {
        const int2 tk = uniform_load(&a.tasks[t]);
        refine_one(a, tk, uniform_load(&a.blocks[tk.x]), L);
    }
}

// Multi-list form (serving): one persistent grid over the frames of several
// streams.  Each list carries its own blocks, workgroup-major task assignment
// and outputs; workgroup g runs, list after list, the tasks of virtual
// workgroup (g + rot) mod grid, so small lists are spread over the grid.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_refine_multi(RefineArgs base, const RefineList *__restrict__ lists, int nLists, int grid)
{
    __shared__ RefineLds L;
    lds_init(L, base);
    for (int l = 0; l < nLists; l++) {
        const RefineList R = uniform_load(&lists[l]);
        RefineArgs a = base;
        a.blocks = R.blocks;
        a.candCost = R.candCost;
        a.best = R.best;
        a.bestKey = R.bestKey;
        a.done = R.done;
        a.mvField = R.mvField;
        a.mvStride = R.mvStride;
        int v = (int)blockIdx.x + R.rot;
        if (v >= grid) v -= grid;
        const int e = uniform(R.wgBegin[v + 1]);
        for (int t = uniform(R.wgBegin[v]); t < e; t++) {
            const int2 tk = uniform_load(&R.tasks[t]);
            refine_one(a, tk, uniform_load(&R.blocks[tk.x]), L);
        }
    }
}

// Plain form: one launch per block shape, one workgroup per task.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_refine_shape(RefineArgs a)
{
    __shared__ RefineLds L;
    lds_init(L, a);
    const int t = a.taskBegin + (int)blockIdx.x;
    if (t < a.taskEnd) {
        const int2 tk = uniform_load(&a.tasks[t]);
        refine_one(a, tk, uniform_load(&a.blocks[tk.x]), L);
    }
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
    int persistFactor = 16;            // persistent workgroups per CU (registers allow 4 waves per SIMD)
    int nregions = 1;                  // spatial partitions of the batched form (XCD-aligned)

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
    uint64_t stateVersion = 1;         // bumped by every picture, prediction and block setter
    uint64_t validKey = 0;             // (state, prediction argument, CTU size) last validated
    int lastNcand = 0;                 // ncand of the last refinement

    // per-shape work queues
    int tile[kNumShapes];              // candidates per task
    std::vector<int2> tasks;
    int2 *dTasks = nullptr;            // shape-major (plain mode)
    int2 *dTasksW = nullptr;           // workgroup-major (batched mode)
    int *dWgBegin = nullptr;           // task range of each persistent workgroup
    int wgCap = 0;
    int gridW = 0;                     // persistent workgroups of the current assignment
    size_t taskCap = 0;
    int shapeTaskBegin[kNumShapes + 1] = {};
    std::vector<int> shapeOrder;       // shapes, heaviest tasks first
    uint64_t tasksKey = 0;             // (blocksVersion, ncand, tile table) the tasks were built for

    uint32_t *dCandCost = nullptr;
    xg_block_result *dBest = nullptr;
    int16_t *dOff = nullptr;           // candidate offsets of the last configuration
    int16_t hOff[2 * XG_MAX_CAND];
    bool offValid = false;
    int *dDone = nullptr;              // per-block candidate counters
    unsigned long long *dBestKey = nullptr;  // per-block best (cost << 32 | cand)
    void *hStage = nullptr;            // pinned upload staging
    size_t hStageCap = 0;

    // refinement prepared by refine_prepare for a caller that enqueues it itself
    RefineArgs prepared;
    bool preparedValid = false;
    void *pipe = nullptr;              // motion-search pipeline state (xg_pipeline.hip)

    // graph of the batched stages
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    uint64_t graphKey = 0;

    // timing
    bool timing = false;
    std::vector<hipEvent_t> evPool;
    std::vector<std::pair<int, int>> evPending[kTimerStages];  // per stage: (start, stop) indices
    int evNext = 0;
    double accMs[kTimerStages] = {};
    int accN[kTimerStages] = {};
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
    v->blocksVersion = s->blocksVersion;
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
            (void)hipHostFree(s->hStage);
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

// Estimated cost of a task in row-passes of the refinement kernel: luma and
// chroma passes (units of W/2 lanes, 64 lanes per pass) times the rows each
// filters, plus a fixed per-task share (setup, staging latency, costs, MC).
// Measured on MI355X at 1080p: ~0.2 us per row-pass and ~8 us per task of
// workgroup time, hence the fixed share of 42.
static long task_cost(const xg_block &B, int nc)
{
    const long lp = (nc * (long)B.w / 2 + 63) / 64, cp = (2L * nc * std::max(1, B.w / 4) + 63) / 64;
    return lp * (B.h + 7) + cp * (B.h / 2 + 3) + 42;
}

// Assignment of the tasks to the persistent workgroups: the picture area is
// split into nregions horizontal bands of equal estimated cost, band r goes to
// the workgroups g with g % nregions == r (workgroup g runs on XCD g % 8, so a
// band's reference rows stay in one XCD's L2), and inside a band each task,
// heaviest first, goes to the least loaded workgroup.
static void assign_tasks(const xg_block *blocks, int nregions, const std::vector<int2> &tasks, int G,
                         std::vector<int2> &out, std::vector<int> &begin)
{
    const int R = std::max(1, std::min(nregions, G));
    const size_t n = tasks.size();
    std::vector<long> cost(n);
    long total = 0;
    for (size_t i = 0; i < n; i++) {
        cost[i] = task_cost(blocks[tasks[i].x], tasks[i].y >> 16);
        total += cost[i];
    }
    // bands: tasks ordered by picture position, cut at equal cost
    std::vector<int> band(n, 0);
    if (R > 1) {
        std::vector<size_t> pos(n);
        for (size_t i = 0; i < n; i++) pos[i] = i;
        std::stable_sort(pos.begin(), pos.end(), [&](size_t x, size_t y) {
            const xg_block &A = blocks[tasks[x].x], &B = blocks[tasks[y].x];
            return A.pic != B.pic ? A.pic < B.pic : A.y != B.y ? A.y < B.y : A.x < B.x;
        });
        long acc = 0;
        for (size_t k = 0; k < n; k++) {
            band[pos[k]] = (int)std::min<long>(R - 1, acc * R / std::max(1L, total));
            acc += cost[pos[k]];
        }
    }
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    std::vector<std::vector<int>> lists(G);
    std::vector<std::priority_queue<std::pair<long, int>, std::vector<std::pair<long, int>>, std::greater<>>> heaps(R);
    for (int g = 0; g < G; g++) heaps[g % R].push({0, g});
    for (size_t i : order) {
        auto &h = heaps[band[i]];
        auto top = h.top();
        h.pop();
        lists[top.second].push_back((int)i);
        h.push({top.first + cost[i], top.second});
    }
    out.clear();
    out.reserve(n);
    begin.assign(G + 1, 0);
    for (int g = 0; g < G; g++) {
        begin[g] = (int)out.size();
        for (int i : lists[g]) out.push_back(tasks[i]);
    }
    begin[G] = (int)out.size();
}

// Tasks of `n` blocks (per-shape candidates per task from the tile table),
// shapes heaviest first: shape-major list and each shape's first task.
static void make_tasks(const int *tile, const xg_block *blocks, int n, int ncand, std::vector<int2> &tasks,
                       int st[kNumShapes + 1], std::vector<int> &order)
{
    std::vector<std::vector<int>> per(kNumShapes);
    for (int b = 0; b < n; b++) per[shape_of(blocks[b].w, blocks[b].h)].push_back(b);
    order.resize(kNumShapes);
    for (int i = 0; i < kNumShapes; i++) order[i] = i;
    auto weight = [&](int sh) {
        const int w = 4 << (sh / 5), h = 4 << (sh % 5);
        return (long)w * h * std::min(tile[sh], ncand);
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return weight(x) > weight(y); });
    tasks.clear();
    for (int oi = 0; oi < kNumShapes; oi++) {
        const int sh = order[oi];
        st[oi] = (int)tasks.size();
        const int cpt = std::max(1, std::min(tile[sh], ncand));
        for (int b : per[sh])
            for (int c0 = 0; c0 < ncand; c0 += cpt)
                tasks.push_back(make_int2(b, c0 | (std::min(cpt, ncand - c0) << 16)));
    }
    st[kNumShapes] = (int)tasks.size();
}

static bool build_tasks(xg_session *s, int ncand)
{
    const int G = std::max(1, s->numCUs * s->persistFactor);
    uint64_t key = mix(mix(s->blocksVersion, (uint64_t)ncand), 0);
    for (int i = 0; i < kNumShapes; i++) key = mix(key, (uint64_t)s->tile[i]);
    key = mix(key, ((uint64_t)s->nregions << 32) | (uint64_t)G);
    if (key == s->tasksKey) return true;
    int st[kNumShapes + 1];
    make_tasks(s->tile, s->blocks.data(), s->nblocks, ncand, s->tasks, st, s->shapeOrder);
    memcpy(s->shapeTaskBegin, st, sizeof(st));
    std::vector<int2> tasksW;
    std::vector<int> wgBegin;
    assign_tasks(s->blocks.data(), s->nregions, s->tasks, G, tasksW, wgBegin);
    if (s->tasks.size() > s->taskCap || G + 1 > s->wgCap) {
        if (!XG_HIP(hipStreamSynchronize(s->stream))) return false;
        (void)hipFree(s->dTasks);
        (void)hipFree(s->dTasksW);
        (void)hipFree(s->dWgBegin);
        s->dTasks = s->dTasksW = nullptr;
        s->dWgBegin = nullptr;
        s->taskCap = 0;
        s->wgCap = 0;
        const size_t cap = std::max<size_t>(s->tasks.size(), 1024);
        const int wcap = std::max(G + 1, 4096);
        if (!XG_HIP(hipMalloc((void **)&s->dTasks, cap * sizeof(int2))) ||
            !XG_HIP(hipMalloc((void **)&s->dTasksW, cap * sizeof(int2))) ||
            !XG_HIP(hipMalloc((void **)&s->dWgBegin, wcap * sizeof(int))))
            return false;
        s->taskCap = cap;
        s->wgCap = wcap;
    }
    if (!s->tasks.empty() && !stage_upload(s, s->dTasks, s->tasks.data(), s->tasks.size() * sizeof(int2))) return false;
    if (!tasksW.empty() && !stage_upload(s, s->dTasksW, tasksW.data(), tasksW.size() * sizeof(int2))) return false;
    if (!stage_upload(s, s->dWgBegin, wgBegin.data(), wgBegin.size() * sizeof(int))) return false;
    if (!XG_HIP(hipStreamSynchronize(s->stream))) return false;
    s->gridW = G;
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
    for (int st = 0; st < kTimerStages; st++) {
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

bool session_timer_begin(xg_session *s, int stage, int ev[2])
{
    if (!s->timing || stage < 0 || stage >= kTimerStages) return false;
    ev[0] = ev_get(s);
    ev[1] = ev_get(s);
    if (ev[0] < 0 || ev[1] < 0) return false;
    return hipEventRecord(s->evPool[ev[0]], s->stream) == hipSuccess;
}

void session_timer_end(xg_session *s, int stage, const int ev[2])
{
    if (hipEventRecord(s->evPool[ev[1]], s->stream) == hipSuccess) s->evPending[stage].push_back({ev[0], ev[1]});
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
                    ((B.y + B.h) >> sub) > pl.height || (pl.stride & 1) || ((uintptr_t)pl.data & 3)) {
                    set_error(XG_ERR_ARG, "prediction plane %d does not cover block %d (or is not dword aligned)", p,
                              i);
                    return false;
                }
            }
        }
    return true;
}

// Prediction picture of every block's slot (when given) and block validation.
// Both depend only on the session state, the prediction argument and the CTU
// size, so a call with all three as at the last successful check skips the
// per-block loops (they cost about as much host time as a 1080p frame's kernel).
static bool prepare_blocks(xg_session *s, const xg_picture *pred, int ctu)
{
    uint64_t key = mix(mix(s->stateVersion, (uint64_t)ctu), pred ? 1 : 0);
    if (pred) {
        const unsigned char *bytes = reinterpret_cast<const unsigned char *>(pred);
        for (size_t i = 0; i < sizeof(xg_picture); i++) key = mix(key, bytes[i]);
    }
    if (key == s->validKey) return true;
    s->validKey = 0;
    if (pred)   // one prediction picture for every current-picture slot in use
        for (int i = 0; i < s->nblocks; i++) {
            s->preds[s->blocks[i].pic] = *pred;
            s->predSet[s->blocks[i].pic] = true;
        }
    if (!validate_blocks(s, true, ctu)) return false;
    s->validKey = key;
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
    const int grid = s->gridW;
    ra.wgBegin = s->dWgBegin;
    ra.tasks = s->dTasksW;
    ra.taskBegin = 0;
    ra.taskEnd = (int)s->tasks.size();
    int e0 = -1, e1 = -1;
    if (timed) {
        e0 = ev_get(s);
        e1 = ev_get(s);
        if (e0 < 0 || e1 < 0) timed = false;
    }
    if (timed) (void)hipEventRecord(s->evPool[e0], s->stream);
    k_refine_persistent<<<grid, 64, 0, s->stream>>>(ra);
    if (timed) {
        (void)hipEventRecord(s->evPool[e1], s->stream);
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
    if (timed) (void)hipEventRecord(s->evPool[e0], s->stream);
    for (int oi = 0; oi < kNumShapes; oi++) {
        const int b0 = s->shapeTaskBegin[oi], b1 = s->shapeTaskBegin[oi + 1];
        if (b1 <= b0) continue;
        ra.wgBegin = nullptr;
        ra.tasks = s->dTasks;
        ra.taskBegin = b0;
        ra.taskEnd = b1;
        k_refine_shape<<<b1 - b0, 64, 0, s->stream>>>(ra);
    }
    if (timed) {
        (void)hipEventRecord(s->evPool[e1], s->stream);
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
              XG_HIP(hipMemset(s->dPics, 0, sizeof(xg_picture) * kPicSlots)) &&      // unset slots: null planes
              XG_HIP(hipMemset(s->dPred, 0, sizeof(xg_picture) * kPicSlots)) &&
              XG_HIP(hipMalloc((void **)&s->dBlocks, sizeof(xg_block) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dCandCost, sizeof(uint32_t) * (size_t)maxBlocks * XG_MAX_CAND)) &&
              XG_HIP(hipMalloc((void **)&s->dBest, sizeof(xg_block_result) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dDone, sizeof(int) * (size_t)maxBlocks)) &&
              XG_HIP(hipMalloc((void **)&s->dBestKey, sizeof(unsigned long long) * (size_t)maxBlocks)) &&
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
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    fdct_release(s);
    recon_release(s);
    dmvr_release(s);
    lookahead_release(s);
    mctf_release(s);
    bipred_release(s);
    quant_release(s);
    pipe_release(s);
    if (s->graphExec) (void)hipGraphExecDestroy(s->graphExec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (auto e : s->evPool) (void)hipEventDestroy(e);
    (void)hipFree(s->dPics);
    (void)hipFree(s->dPred);
    (void)hipFree(s->dBlocks);
    (void)hipFree(s->dTasks);
    (void)hipFree(s->dTasksW);
    (void)hipFree(s->dWgBegin);
    (void)hipFree(s->dCandCost);
    (void)hipFree(s->dBest);
    (void)hipFree(s->dDone);
    (void)hipFree(s->dBestKey);
    (void)hipFree(s->dOff);
    if (s->hStage) (void)hipHostFree(s->hStage);
    if (s->stream) (void)hipStreamDestroy(s->stream);
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
    s->stateVersion++;
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
    s->stateVersion++;
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
    s->stateVersion++;
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
    if (!prepare_blocks(s, pred, cfg->ctuSize)) return XG_ERR_ARG;
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
        if (s->graphExec) (void)hipGraphExecDestroy(s->graphExec);
        if (s->graph) (void)hipGraphDestroy(s->graph);
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
    if (timed) (void)hipEventRecord(s->evPool[e0], s->stream);
    if (!XG_HIP(hipGraphLaunch(s->graphExec, s->stream))) return XG_ERR_HIP;
    if (timed) {
        (void)hipEventRecord(s->evPool[e1], s->stream);
        s->evPending[1].push_back({e0, e1});
    }
    return XG_OK;
}

}  // extern "C"

namespace xg {
int32_t refine_prepare(xg_session *s, const xg_refine_cfg *cfg, const xg_picture *pred)
{
    s->preparedValid = false;
    if (!validate_cfg(cfg)) return XG_ERR_ARG;
    if (s->nblocks == 0) return XG_ERR_STATE;
    if (!prepare_blocks(s, pred, cfg->ctuSize)) return XG_ERR_ARG;
    if (!build_tasks(s, cfg->ncand)) return XG_ERR_HIP;
    if (!fill_args(s, cfg, s->prepared, true)) return XG_ERR_HIP;
    s->lastNcand = cfg->ncand;
    s->preparedValid = true;
    return XG_OK;
}

bool refine_enqueue_prepared(xg_session *s)
{
    return s->preparedValid && enqueue_batched(s, s->prepared, false);
}

int refine_grid(xg_session *s) { return std::max(1, s->numCUs * s->persistFactor); }

void refine_assign(xg_session *s, const xg_block *blocks, int n, int ncand, std::vector<int2> &tasksW,
                   std::vector<int> &wgBegin)
{
    std::vector<int2> tasks;
    std::vector<int> order;
    int st[kNumShapes + 1];
    make_tasks(s->tile, blocks, n, ncand, tasks, st, order);
    assign_tasks(blocks, s->nregions, tasks, refine_grid(s), tasksW, wgBegin);
}

int32_t refine_multi_enqueue(xg_session *s, const xg_refine_cfg *cfg, const RefineList *dLists, int nLists)
{
    if (!validate_cfg(cfg)) return XG_ERR_ARG;
    if (nLists <= 0) return XG_OK;
    RefineArgs base;
    if (!fill_args(s, cfg, base, true)) return XG_ERR_HIP;
    int ev[2];
    const bool timed = session_timer_begin(s, 0, ev);
    k_refine_multi<<<refine_grid(s), 64, 0, s->stream>>>(base, dLists, nLists, refine_grid(s));
    if (timed) session_timer_end(s, 0, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

uint64_t refine_prepared_key(xg_session *s)
{
    const RefineArgs &a = s->prepared;
    uint64_t k = mix(s->tasksKey, ((uint64_t)s->gridW << 32) | (uint64_t)s->tasks.size());
    for (const void *ptr : {(const void *)a.pics, (const void *)a.blocks, (const void *)s->dTasksW,
                            (const void *)s->dWgBegin, (const void *)a.candCost, (const void *)a.best,
                            (const void *)a.bestKey, (const void *)a.done, (const void *)a.off,
                            (const void *)a.predTab})
        k = mix(k, (uint64_t)(uintptr_t)ptr);
    k = mix(k, ((uint64_t)(uint32_t)a.ncand << 32) | (uint32_t)a.ctuSize);
    k = mix(k, ((uint64_t)a.lambda << 32) | ((uint64_t)(uint32_t)a.useChroma << 1) | (uint64_t)(a.doMc != 0));
    k = mix(k, ((uint64_t)(uint16_t)a.offMinX << 48) | ((uint64_t)(uint16_t)a.offMaxX << 32) |
                   ((uint64_t)(uint16_t)a.offMinY << 16) | (uint64_t)(uint16_t)a.offMaxY);
    return mix(k, (uint64_t)s->preparedValid);
}

void *&session_pipe(xg_session *s) { return s->pipe; }
}  // namespace xg

extern "C" {

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
    for (int st = 0; st < kTimerStages; st++) {
        s->accMs[st] = 0;
        s->accN[st] = 0;
    }
    return XG_OK;
}

// stage 0: refinement kernel(s) of the eager modes; 1: whole graph replay;
// 2: forward-transform kernel; 3: integer-search kernel.
int32_t xg_session_kernel_time(xg_session *s, int32_t stage, float *avgMs, int32_t *count)
{
    if (!s || stage < 0 || stage >= kTimerStages) return XG_ERR_ARG;
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
        if (!XG_HIP(hipEventElapsedTime(&ms, e0, e1))) return 0.f;
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
        // more bands keep each band's reference rows in one XCD's L2 (8 bands fetch
        // about a quarter of what 2 do), which frame time alone does not show: a
        // schedule with more bands wins unless it is over 1% slower
        int bestR = s->nregions, bestF = s->persistFactor;
        for (int r : {8, 4, 2, 1})
            for (int f : {16, 14, 12, 10, 8}) {
                s->nregions = r;
                s->persistFactor = f;
                const float ms = time_frame(s, cfg, reps, e0, e1, &ok);
                if (!ok) { rc = XG_ERR_HIP; break; }
                if (ms < bestMs * (r < bestR ? 0.99f : 1.0f)) { bestMs = ms; bestR = r; bestF = f; }
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
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
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

// xg_amvp.hip -- AMVP predictor choice (Xin266SelectMvp) as a batched stage.
//
// Per block, each AMVP candidate is rounded to the AMVR precision and back
// (the end of Xin266FillAmvpCand, h266_get_neighbour_mv.c:2549-2558), its luma
// prediction is formed as Xin266LumaInterPred forms it
// (h266_inter_prediction.c:302-354: MV clipped to the CU window, table
// dispatch of the 8-tap filter) and compared with the block by SAD; the
// first strictly lowest wins (h266_analyze_inter_cu.c:97-126).
//
// Mapping: eight lanes per block, no LDS and no barriers.  Lane q of a
// block's eight costs candidate q >> 2 on its share of the block's 4-column
// strips (blocks 4 or 8 wide are also split into row segments, so the four
// lanes of a candidate stay busy).  A strip walks down its reference rows: each row gives the four first-pass values of the strip (8 taps as
// int16 pairs on v_dot2), which pair with the previous row's into the
// register ring the vertical pass reads, and from the eighth row on each row
// completes one output row of four predicted samples, compared with the
// current block at once.  The strip SADs add over the four lanes of a
// candidate by quad permutes, and the two candidates' lanes exchange theirs.
// The separable two-pass form with the phase-0 tap (64 at the centre) gives
// exactly the copy, horizontal-only and vertical-only results the reference's
// table dispatch produces, so every candidate runs the same code.
//
// The candidate lists either come from the caller (xg_amvp_select) or are
// derived from an MV field of 4x4 units (xg_amvp_from_field, and the serving
// stage k_amvp_multi, which reads each stream's resident field).
#include <algorithm>

#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

constexpr int kAmvpMaxSide = 64;
constexpr int kAmvpLanes = 8;                          // lanes per block
constexpr int kAmvpWg = 256;
constexpr int kAmvpPerWg = kAmvpWg / kAmvpLanes;       // blocks per workgroup
constexpr int kAmvpMaxLists = 32;                      // lists of one k_amvp_multi launch

// amvrPrecision (h266_entropy_write.c:120-123): log2 units per 4 pel by imvIdx.
__device__ __forceinline__ int amvr_prec(int imvIdx)
{
    return imvIdx == 1 ? 2 : imvIdx == 2 ? 0 : imvIdx == 3 ? 3 : 4;
}

// Xin266ChangeMv32Prec (h266_get_neighbour_mv.c:212-243): shift left when the
// precision rises, round half toward zero when it drops.
__device__ __forceinline__ int change_prec(int v, int inPrec, int outPrec)
{
    const int shift = outPrec - inPrec;
    if (shift >= 0) return v * (1 << shift);
    const int s = -shift, ofs = 1 << (s - 1);
    return v >= 0 ? (v + ofs - 1) >> s : (v + ofs) >> s;
}

// The 8 taps of a filter row, times 64, as four int16 pairs.
__device__ __forceinline__ void taps_x64(const int8_t *t, uint32_t *p)
{
#pragma unroll
    for (int m = 0; m < 4; m++) p[m] = pack16(64 * (int)t[2 * m], 64 * (int)t[2 * m + 1]);
}

// The current samples (c, c + 1) and (c + 2, c + 3) of a row as two dwords.
__device__ __forceinline__ void cur_pairs(const gword *c, bool aligned, uint32_t &p01, uint32_t &p23)
{
    if (aligned) {
        p01 = *(gdword *)c;
        p23 = *(gdword *)(c + 2);
    } else {
        p01 = pack16(c[0], c[1]);
        p23 = pack16(c[2], c[3]);
    }
}

// Prediction SAD of one candidate over the 4-column strips s0, s0 + ds, ...
// (< ns) of rows [0, n) of a block: rbase is the reference sample three rows
// and three columns above-left of the candidate's first predicted sample
// (the clipped integer position), cbase the block's first current sample.
//
// Row r of a strip at column cs needs reference samples cs .. cs + 10.  They
// are read as the five dwords from cs rounded down to an even sample plus
// either the dword ending at cs + 10 (odd start) or the sample cs + 10 alone
// (even start), so no load leaves [cs - 1, cs + 10].  With od the start's
// parity, B_k = alignbit(A_{k+1}, A_k, 16 od) is the sample pair (cs + 2k,
// cs + 2k + 1) and O_k = alignbit(B_{k+1}, B_k, 16) the pair (cs + 2k + 1,
// cs + 2k + 2); output column j of the strip filters B_{j/2} .. for even j
// and O_{(j-1)/2} .. for odd j.
//
// Both passes use the taps times 64 (pack_mid16: the first-pass value of two
// rows in one dword; clip_hi16: the clipped second pass in the upper half);
// the first pass is the offset-free form (sum + 32) >> 2, whose vertical sum
// is >> 10 alone (the taps of each filter add to 64).  Rows are walked with
// one ring of eight row pairs per column and loaded four rows ahead; rows
// past the strip re-read its last row.
__device__ uint32_t strip_sad(const gword *rbase, intptr_t rs, const gword *cbase, intptr_t cst, bool curAligned,
                              int n, int s0, int ds, int ns, const uint32_t *ht, const uint32_t *vt)
{
    const int nrows = n + 7;
    uint32_t sad = 0;
    for (int s = s0; s < ns; s += ds) {
        const gword *lp = rbase + 4 * s;               // next reference row to load
        const gword *lc = cbase + 4 * s;               // next current row to load
        int prev[4] = {0, 0, 0, 0};
        uint32_t ring[8][4];                           // ring[p & 7][j]: rows (p, p + 1) of column j
        for (int r0 = 0; r0 < nrows; r0 += 8) {
            uint32_t A[4][6], sh[4], C01[4], C23[4];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if ((k & 3) == 0) {
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int r = r0 + k + i;
                        const uintptr_t a = (uintptr_t)lp;
                        const gdword *q = (const gdword *)(a & ~(uintptr_t)3);
#pragma unroll
                        for (int m = 0; m < 5; m++) A[i][m] = q[m];
                        const uint32_t tail = *(const gdword *)((a + 18) & ~(uintptr_t)3);
                        const uint32_t last = lp[10];
                        A[i][5] = (a & 2) ? tail : last;
                        sh[i] = (uint32_t)(a & 2) << 3;        // 16 for an odd start
                        if (r + 1 < nrows) lp += rs;
                        cur_pairs(lc, curAligned, C01[i], C23[i]);
                        if (r >= 7 && r + 1 < nrows) lc += cst;
                    }
                }
                const int r = r0 + k, i = k & 3;
                if (r >= nrows) break;
                uint32_t B[6], O[5];
#pragma unroll
                for (int m = 0; m < 5; m++) B[m] = align16(A[i][m + 1], A[i][m], sh[i]);
                B[5] = align16(A[i][5], A[i][5], sh[i]);
#pragma unroll
                for (int m = 0; m < 5; m++) O[m] = align16(B[m + 1], B[m], 16);
                // first pass of the four columns, the chains interleaved
                int hs[4];
#pragma unroll
                for (int j = 0; j < 4; j++) hs[j] = dot2_first((j & 1) ? O[j >> 1] : B[j >> 1], ht[0], kHRound64);
#pragma unroll
                for (int m = 1; m < 4; m++)
#pragma unroll
                    for (int j = 0; j < 4; j++) hs[j] = dot2((j & 1) ? O[(j >> 1) + m] : B[(j >> 1) + m], ht[m], hs[j]);
                if (r >= 1) {
#pragma unroll
                    for (int j = 0; j < 4; j++) ring[(k + 7) & 7][j] = pack_mid16(prev[j], hs[j]);
                }
#pragma unroll
                for (int j = 0; j < 4; j++) prev[j] = hs[j];
                if (r >= 7) {                          // output row r - 7: pairs from rows r - 7, r - 5, r - 3, r - 1
                    int v[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) v[j] = dot2_first(ring[(k + 1) & 7][j], vt[0], 0);
#pragma unroll
                    for (int m = 1; m < 4; m++)
#pragma unroll
                        for (int j = 0; j < 4; j++) v[j] = dot2(ring[(k + 1 + 2 * m) & 7][j], vt[m], v[j]);
                    sad = __builtin_amdgcn_sad_u16(pack_hi16(clip_hi16(v[0]), clip_hi16(v[1])), C01[i], sad);
                    sad = __builtin_amdgcn_sad_u16(pack_hi16(clip_hi16(v[2]), clip_hi16(v[3])), C23[i], sad);
                }
            }
        }
    }
    return sad;
}

// The lane's share of one candidate's prediction SAD for the w x h block at
// (x, y) (mvx, mvy in 1/16 pel, before the CU-window clip): lane g of the
// candidate's four.  A block of ns 4-column strips is split into
// 4 / min(ns, 4) row segments so narrow blocks keep all four lanes busy: each
// lane takes the strips g % L, g % L + L, ... (L = 4 / segments) of segment
// g / L.
__device__ uint32_t candidate_sad(const xg_plane &cur, const xg_plane &ref, int x, int y, int w, int h, int mvx,
                                  int mvy, int fi, int ctu, int g)
{
    const int fw = cur.width, fh = cur.height;
    // CU MV window (h266_enc_init.c:1462-1465, int16 as stored there), 1/16 pel
    const int mnx = (int)(int16_t)(-x - ctu - 8) * 16, mny = (int)(int16_t)(-y - ctu - 8) * 16;
    const int mxx = (int)(int16_t)(fw + (ctu - w + 8) - x) * 16, mxy = (int)(int16_t)(fh + (ctu - h + 8) - y) * 16;
    const int ox = x * 16 + min(max(mvx, mnx), mxx), oy = y * 16 + min(max(mvy, mny), mxy);
    uint32_t ht[4], vt[4];
    taps_x64(taps_for<8>(ox & 15, fi), ht);
    taps_x64(taps_for<8>(oy & 15, fi), vt);
    const int ns = w >> 2, segs = ns == 1 ? 4 : ns == 2 ? 2 : 1, L = 4 / segs;
    const int n = h / segs, seg = g / L;
    const intptr_t rs = ref.stride, cst = cur.stride;
    const gword *rbase = (const gword *)ref.data + (intptr_t)((oy >> 4) - 3 + seg * n) * rs + ((ox >> 4) - 3);
    const gword *cbase = (const gword *)cur.data + (intptr_t)(y + seg * n) * cst + x;
    const bool aligned = (((uintptr_t)cbase & 3) | (cst & 1)) == 0;
    return strip_sad(rbase, rs, cbase, cst, aligned, n, g % L, L, ns, ht, vt);
}

// Whether a block can be costed: inside its picture, sides 4..64 and
// multiples of 4, both slots set, the reference padded for the CU window.
__device__ __forceinline__ bool amvp_block_ok(const xg_plane &cur, const xg_plane &ref, int x, int y, int w, int h,
                                              int ctu)
{
    return cur.data && ref.data && w >= 4 && h >= 4 && w <= kAmvpMaxSide && h <= kAmvpMaxSide && !(w & 3) &&
           !(h & 3) && x >= 0 && y >= 0 && x + w <= cur.width && y + h <= cur.height && ref.pad >= ctu + 12 &&
           ref.width == cur.width && ref.height == cur.height;
}

// Xin266SelectMvp for one block, by the eight lanes of its group (every lane
// of the wave calls it: the sums cross lanes).  Returns the winner index
// (-1: rejected) with its SAD and rounded MV in every lane of the group.
// cost false: a single candidate is taken without costing it.
__device__ int select_mvp(const xg_picture *pics, bool live, int pic, int refSlot, int x, int y, int w, int h,
                          int ncand, int c0x, int c0y, int c1x, int c1y, int imvIdx, int ctu, bool cost,
                          uint32_t &sadOut, int &mvx, int &mvy)
{
    const bool slots = live && pic >= 0 && pic < kSessionPicSlots && refSlot >= 0 && refSlot < kSessionPicSlots;
    xg_plane cur{}, ref{};
    if (slots) {
        cur = pics[pic].plane[0];
        ref = pics[refSlot].plane[0];
    }
    const bool ok = slots && amvp_block_ok(cur, ref, x, y, w, h, ctu);
    const int prec = amvr_prec(imvIdx), fi = imvIdx == 3 ? 1 : 0;
    const int m0x = change_prec(change_prec(c0x, 6, prec), prec, 6);
    const int m0y = change_prec(change_prec(c0y, 6, prec), prec, 6);
    const int m1x = change_prec(change_prec(c1x, 6, prec), prec, 6);
    const int m1y = change_prec(change_prec(c1y, 6, prec), prec, 6);
    // two candidates equal after the rounding predict the same samples: the first wins
    // their tie, so only it is costed (and not at all when no SAD is asked for)
    const int nc = min(max(ncand, 1), 2) == 2 && (m0x != m1x || m0y != m1y) ? 2 : 1;
    const int q = threadIdx.x & (kAmvpLanes - 1), c = q >> 2;
    uint32_t part = 0;
    if (ok && (cost || nc > 1) && c < nc)
        part = candidate_sad(cur, ref, x, y, w, h, c ? m1x : m0x, c ? m1y : m0y, fi, ctu, q & 3);
    part = (uint32_t)group_sum<4>((int)part);          // over the candidate's four strip lanes
    const uint32_t other = (uint32_t)shfl_xor<4>((int)part);
    const uint32_t s0 = c ? other : part, s1 = c ? part : other;
    const int idx = nc > 1 && s1 < s0 ? 1 : 0;         // the first strictly lowest
    if (!ok) {
        sadOut = 0xFFFFFFFFu;
        mvx = mvy = 0;
        return -1;
    }
    sadOut = idx ? s1 : s0;
    mvx = idx ? m1x : m0x;
    mvy = idx ? m1y : m0y;
    return idx;
}

// The AMVP list of Xin266FillAmvpCand (h266_get_neighbour_mv.c:2451-2560) from
// an MV field: every unit inside the picture is an available inter neighbour.
__device__ __forceinline__ bool field_unit(const int2 *field, int stride, int fw, int fh, int px, int py, int2 &mv)
{
    if (px < 0 || py < 0 || px >= fw || py >= fh) return false;
    mv = field[(intptr_t)(py >> 2) * stride + (px >> 2)];
    return true;
}

__device__ int amvp_list(const int2 *field, int stride, int fw, int fh, int x, int y, int w, int h, int *cand)
{
    int2 a = make_int2(0, 0), b = make_int2(0, 0);
    const bool hasA = field_unit(field, stride, fw, fh, x - 1, y + h, a) ||       // A0 below-left
                      field_unit(field, stride, fw, fh, x - 1, y + h - 1, a);     // A1 left
    const bool hasB = field_unit(field, stride, fw, fh, x + w, y - 1, b) ||       // B0 above-right
                      field_unit(field, stride, fw, fh, x + w - 1, y - 1, b) ||   // B1 above
                      field_unit(field, stride, fw, fh, x - 1, y - 1, b);         // B2 above-left
    // the list so far: A then B, a duplicate pair reduced to one
    int2 l0 = hasA ? a : b, l1 = b;
    int n = (int)hasA + (int)hasB;
    if (n == 2 && a.x == b.x && a.y == b.y) n = 1;
    // a zero MV joins an empty list or one whose single entry is non-zero
    if (n == 0) l0 = make_int2(0, 0), n = 1;
    else if (n == 1 && (l0.x != 0 || l0.y != 0)) l1 = make_int2(0, 0), n = 2;
    const bool two = n == 2 && (l0.x != l1.x || l0.y != l1.y);
    cand[0] = l0.x;
    cand[1] = l0.y;
    cand[2] = two ? l1.x : 0;
    cand[3] = two ? l1.y : 0;
    return two ? 2 : 1;
}

__global__ void __launch_bounds__(kAmvpWg) k_amvp_select(const xg_picture *pics,
                                                         const xg_amvp_block *__restrict__ blocks, int n, int imvIdx,
                                                         int ctu, xg_amvp_result *__restrict__ out)
{
    const int b = (int)blockIdx.x * kAmvpPerWg + (int)(threadIdx.x / kAmvpLanes);
    const bool live = b < n;
    xg_amvp_block B{};
    if (live) B = blocks[b];
    uint32_t sad;
    int mvx, mvy;
    const int idx = select_mvp(pics, live, B.pic, B.ref, B.x, B.y, B.w, B.h, B.ncand, B.cand[0][0], B.cand[0][1],
                               B.cand[1][0], B.cand[1][1], imvIdx, ctu, true, sad, mvx, mvy);
    if (live && (threadIdx.x & (kAmvpLanes - 1)) == 0) {
        xg_amvp_result r;
        r.mvpIdx = idx;
        r.sad = sad;
        r.mv[0] = mvx;
        r.mv[1] = mvy;
        r.meMv[0] = (int16_t)change_prec(mvx, 6, 4);      // Xin266ChangeMv2Me: 1/16 -> quarter pel
        r.meMv[1] = (int16_t)change_prec(mvy, 6, 4);
        out[b] = r;
    }
}

__global__ void k_amvp_from_field(const int2 *field, int stride, int fw, int fh, const xg_block *__restrict__ blocks,
                                  int n, xg_amvp_block *__restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const xg_block B = blocks[b];
    int cand[4];
    xg_amvp_block o;
    o.x = B.x;
    o.y = B.y;
    o.w = B.w;
    o.h = B.h;
    o.pic = B.pic;
    o.ref = B.ref;
    o.ncand = amvp_list(field, stride, fw, fh, B.x, B.y, B.w, B.h, cand);
    o.cand[0][0] = cand[0];
    o.cand[0][1] = cand[1];
    o.cand[1][0] = cand[2];
    o.cand[1][1] = cand[3];
    out[b] = o;
}

// Serving form: the blocks of several streams' frames in one launch (block i
// belongs to the last list whose `blockBegin` is <= i).  Each block takes its
// frame's slots and, as predictor, the quarter-pel AMVR choice among the
// candidates of its stream's MV field (the final MVs of the stream's previous
// frame; zero before the first); slots and predictor are written back.
__global__ void __launch_bounds__(kAmvpWg) k_amvp_multi(const xg_picture *pics, const SearchList *__restrict__ lists,
                                                        int nLists, int nBlocks, int ctu)
{
    __shared__ int begins[kAmvpMaxLists];
    if ((int)threadIdx.x < nLists) begins[threadIdx.x] = lists[threadIdx.x].blockBegin;
    __syncthreads();
    const int gb = (int)blockIdx.x * kAmvpPerWg + (int)(threadIdx.x / kAmvpLanes);
    const bool live = gb < nBlocks;
    int l = 0;
    for (int i = 1; i < nLists; i++) l += begins[i] <= gb;
    const SearchList &S = lists[l];
    const int b = gb - begins[l];
    xg_block B{};
    int cand[4] = {0, 0, 0, 0}, nc = 1;
    if (live) {
        B = S.blocks[b];
        const xg_plane &cur = pics[S.curSlot].plane[0];
        nc = amvp_list(S.mvField, S.mvStride, cur.width, cur.height, B.x, B.y, B.w, B.h, cand);
    }
    uint32_t sad;
    int mvx, mvy;
    const int idx = select_mvp(pics, live, S.curSlot, S.refSlot, B.x, B.y, B.w, B.h, nc, cand[0], cand[1], cand[2],
                               cand[3], 0, ctu, false, sad, mvx, mvy);
    if (live && (threadIdx.x & (kAmvpLanes - 1)) == 0) {
        B.pic = S.curSlot;
        B.ref = S.refSlot;
        B.mvpx = idx >= 0 ? mvx : 0;
        B.mvpy = idx >= 0 ? mvy : 0;
        S.blocks[b] = B;
    }
}

int32_t amvp_multi_enqueue(hipStream_t st, const xg_picture *dPics, const SearchList *dLists, int nLists, int nBlocks,
                           int ctu)
{
    if (nLists <= 0 || nBlocks <= 0) return XG_OK;
    if (nLists > kAmvpMaxLists) return XG_ERR_ARG;
    k_amvp_multi<<<(nBlocks + kAmvpPerWg - 1) / kAmvpPerWg, kAmvpWg, 0, st>>>(dPics, dLists, nLists, nBlocks, ctu);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

}  // namespace xg

using namespace xg;

extern "C" {

int32_t xg_amvp_select(xg_session *s, const xg_amvp_block *blocks, int32_t n, int32_t imvIdx, int32_t ctuSize,
                       xg_amvp_result *out)
{
    SessionView v;
    if (!s || !session_view(s, &v) || n < 0 || (n > 0 && (!blocks || !out)) || imvIdx < 0 || imvIdx > 3 ||
        ctuSize < 8 || ctuSize > 128) {
        set_error(XG_ERR_ARG, "xg_amvp_select: bad arguments (session, arrays, imvIdx 0..3, ctuSize 8..128)");
        return XG_ERR_ARG;
    }
    if (n == 0) return XG_OK;
    k_amvp_select<<<(n + kAmvpPerWg - 1) / kAmvpPerWg, kAmvpWg, 0, v.stream>>>(v.dPics, blocks, n, imvIdx, ctuSize,
                                                                               out);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

int32_t xg_amvp_from_field(xg_session *s, const int32_t *field, int32_t fieldStride, int32_t width, int32_t height,
                           const xg_block *blocks, int32_t n, xg_amvp_block *out)
{
    SessionView v;
    if (!s || !session_view(s, &v) || n < 0 || (n > 0 && (!blocks || !out || !field)) || width < 4 || height < 4 ||
        fieldStride < (width + 3) / 4) {
        set_error(XG_ERR_ARG, "xg_amvp_from_field: bad arguments (session, arrays, picture size, field stride)");
        return XG_ERR_ARG;
    }
    if (n == 0) return XG_OK;
    k_amvp_from_field<<<(n + 255) / 256, 256, 0, v.stream>>>((const int2 *)field, fieldStride, width, height, blocks,
                                                            n, out);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

}  // extern "C"

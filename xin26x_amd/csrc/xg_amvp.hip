// xg_amvp.hip -- AMVP predictor choice (Xin266SelectMvp) as a batched stage.
//
// Per block, each AMVP candidate is rounded to the AMVR precision and back
// (the end of Xin266FillAmvpCand, h266_get_neighbour_mv.c:2549-2558), its luma
// prediction is formed as Xin266LumaInterPred forms it
// (h266_inter_prediction.c:302-354: MV clipped to the CU window, table
// dispatch of the 8-tap filter) and compared with the block by SAD; the
// first strictly lowest wins (h266_analyze_inter_cu.c:97-126).
//
// Mapping: one 64-lane workgroup per block, both candidates at once, in
// sub-blocks of at most 16 x 16.  A candidate's reference window, (H + 7)
// rows of W + 7 samples from three rows / columns above-left of the clipped
// integer position, is staged in LDS; the first filter pass (int16, the
// standard's intermediate) goes to LDS as well, and the second pass, the
// clip and the absolute difference with the current block are one loop whose
// sums reduce across the wave.  The separable two-pass form with the
// phase-0 tap (64 at the centre) gives exactly the copy, horizontal-only and
// vertical-only results the reference's table dispatch produces, so every
// candidate runs the same code.
//
// The candidate lists either come from the caller (xg_amvp_select) or are
// derived from an MV field of 4x4 units (xg_amvp_from_field, and the serving
// stage k_amvp_multi, which reads each stream's resident field).
#include <algorithm>

#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

constexpr int kAmvpMaxSide = 64;
constexpr int kAmvpSub = 16;                           // blocks are costed in sub-blocks of at most 16 x 16
constexpr int kAmvpPitch = kAmvpSub + 8;               // samples per staged row (W + 7 at most)
constexpr int kAmvpRows = kAmvpSub + 7;

struct AmvpLds {
    uint16_t win[2][kAmvpRows * kAmvpPitch];          // both candidates' windows
    int16_t tmp[2][kAmvpRows * kAmvpSub];             // both candidates' first passes
};

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

// Prediction SADs of nc (1 or 2) candidate MVs (1/16 pel, already rounded)
// for the block at (x, y) of w x h: both candidates are staged, filtered and
// compared in the same loops, sub-block by sub-block.  Every lane returns
// the totals in sad[0..nc).
__device__ void candidate_sads(const xg_plane &cur, const xg_plane &ref, int x, int y, int w, int h, int nc,
                               const int *mv, int fi, int ctu, AmvpLds &L, uint32_t *sad)
{
    const int lane = threadIdx.x;
    const int fw = cur.width, fh = cur.height;
    // CU MV window (h266_enc_init.c:1462-1465, int16 as stored there), 1/16 pel
    const int mnx = (int)(int16_t)(-x - ctu - 8) * 16, mny = (int)(int16_t)(-y - ctu - 8) * 16;
    const int mxx = (int)(int16_t)(fw + (ctu - w + 8) - x) * 16, mxy = (int)(int16_t)(fh + (ctu - h + 8) - y) * 16;
    const uint16_t *src[2];
    const int8_t *hf[2], *vf[2];
    for (int c = 0; c < 2; c++) {
        const int k = c < nc ? c : 0;
        const int ox = x * 16 + min(max(mv[2 * k], mnx), mxx), oy = y * 16 + min(max(mv[2 * k + 1], mny), mxy);
        hf[c] = taps_for<8>(ox & 15, fi);
        vf[c] = taps_for<8>(oy & 15, fi);
        src[c] = ref.data + (intptr_t)((oy >> 4) - 3) * ref.stride + ((ox >> 4) - 3);
    }
    const int sw = min(w, kAmvpSub), sh = min(h, kAmvpSub), ww = sw + 7, hh = sh + 7;
    int acc0 = 0, acc1 = 0;
    for (int sy = 0; sy < h; sy += sh)
        for (int sx = 0; sx < w; sx += sw) {
            __syncthreads();                           // the previous sub-block's readers are done
            for (int i = lane; i < nc * hh * ww; i += 64) {
                const int c = i >= hh * ww, j = i - c * hh * ww, r = j / ww, col = j - r * ww;
                L.win[c][r * kAmvpPitch + col] = src[c][(intptr_t)(sy + r) * ref.stride + sx + col];
            }
            __syncthreads();
            for (int i = lane; i < nc * hh * sw; i += 64) {
                const int c = i >= hh * sw, j = i - c * hh * sw, r = j / sw, col = j - r * sw;
                L.tmp[c][r * kAmvpSub + col] = (int16_t)hpass<8>(&L.win[c][r * kAmvpPitch + col], c ? hf[1] : hf[0]);
            }
            __syncthreads();
            const uint16_t *cb = cur.data + (intptr_t)(y + sy) * cur.stride + x + sx;
            for (int i = lane; i < nc * sh * sw; i += 64) {
                const int c = i >= sh * sw, j = i - c * sh * sw, r = j / sw, col = j - r * sw;
                const int8_t *v = c ? vf[1] : vf[0];
                int s = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) s += (int)v[k] * (int)L.tmp[c][(r + k) * kAmvpSub + col];
                const int d = abs((int)cb[(intptr_t)r * cur.stride + col] - vfinal_pix(s));
                if (c) acc1 += d;
                else acc0 += d;
            }
        }
    sad[0] = (uint32_t)group_sum<64>(acc0);
    if (nc > 1) sad[1] = (uint32_t)group_sum<64>(acc1);
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

// Xin266SelectMvp for one block; returns the winner index (-1: rejected).
// needSad false: a single candidate is taken without costing it.
__device__ int select_mvp(const xg_picture *pics, int pic, int refSlot, int x, int y, int w, int h, int ncand,
                          const int *cand, int imvIdx, int ctu, bool needSad, AmvpLds &L, uint32_t *sadOut, int *mvOut)
{
    const bool slots = pic >= 0 && pic < kSessionPicSlots && refSlot >= 0 && refSlot < kSessionPicSlots;
    const xg_plane cur = slots ? pics[pic].plane[0] : xg_plane{}, ref = slots ? pics[refSlot].plane[0] : xg_plane{};
    if (!slots || !amvp_block_ok(cur, ref, x, y, w, h, ctu)) {
        *sadOut = 0xFFFFFFFFu;
        mvOut[0] = mvOut[1] = 0;
        return -1;
    }
    const int prec = amvr_prec(imvIdx), fi = imvIdx == 3 ? 1 : 0;
    ncand = min(max(ncand, 1), 2);
    int mv[4];
    for (int k = 0; k < 2 * ncand; k++) mv[k] = change_prec(change_prec(cand[k], 6, prec), prec, 6);
    uint32_t sad[2] = {0, 0xFFFFFFFFu};
    if (needSad || ncand > 1) candidate_sads(cur, ref, x, y, w, h, ncand, mv, fi, ctu, L, sad);
    const int idx = ncand > 1 && sad[1] < sad[0] ? 1 : 0;  // the first strictly lowest
    *sadOut = sad[idx];
    mvOut[0] = mv[2 * idx];
    mvOut[1] = mv[2 * idx + 1];
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
    int2 l[3];
    int n = 0;
    if (field_unit(field, stride, fw, fh, x - 1, y + h, l[n]) ||        // A0 below-left
        field_unit(field, stride, fw, fh, x - 1, y + h - 1, l[n]))      // A1 left
        n++;
    if (field_unit(field, stride, fw, fh, x + w, y - 1, l[n]) ||        // B0 above-right
        field_unit(field, stride, fw, fh, x + w - 1, y - 1, l[n]) ||    // B1 above
        field_unit(field, stride, fw, fh, x - 1, y - 1, l[n]))          // B2 above-left
        n++;
    if (n == 2 && l[0].x == l[1].x && l[0].y == l[1].y) n = 1;
    if (n < 1 || (n == 1 && (l[0].x != 0 || l[0].y != 0))) l[n++] = make_int2(0, 0);
    const bool two = n >= 2 && (l[0].x != l[1].x || l[0].y != l[1].y);
    cand[0] = l[0].x;
    cand[1] = l[0].y;
    cand[2] = two ? l[1].x : 0;
    cand[3] = two ? l[1].y : 0;
    return two ? 2 : 1;
}

__global__ void __launch_bounds__(64) k_amvp_select(const xg_picture *pics, const xg_amvp_block *__restrict__ blocks,
                                                    int n, int imvIdx, int ctu, xg_amvp_result *__restrict__ out)
{
    __shared__ AmvpLds L;
    const int b = blockIdx.x;
    if (b >= n) return;
    const xg_amvp_block B = blocks[b];
    const int cand[4] = {B.cand[0][0], B.cand[0][1], B.cand[1][0], B.cand[1][1]};
    uint32_t sad;
    int mv[2];
    const int idx = select_mvp(pics, B.pic, B.ref, B.x, B.y, B.w, B.h, B.ncand, cand, imvIdx, ctu, true, L, &sad, mv);
    if (threadIdx.x == 0) {
        xg_amvp_result r;
        r.mvpIdx = idx;
        r.sad = sad;
        r.mv[0] = mv[0];
        r.mv[1] = mv[1];
        r.meMv[0] = (int16_t)change_prec(mv[0], 6, 4);     // Xin266ChangeMv2Me: 1/16 -> quarter pel
        r.meMv[1] = (int16_t)change_prec(mv[1], 6, 4);
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

// Serving form: the blocks of several streams' frames in one launch (workgroup
// w belongs to the last list whose `blockBegin` is <= w).  Each block takes its
// frame's slots and, as predictor, the quarter-pel AMVR choice among the
// candidates of its stream's MV field (the final MVs of the stream's previous
// frame; zero before the first); slots and predictor are written back.
__global__ void __launch_bounds__(64) k_amvp_multi(const xg_picture *pics, const SearchList *__restrict__ lists,
                                                   int nLists, int ctu)
{
    __shared__ AmvpLds L;
    int l = 0;
    while (l + 1 < nLists && (int)blockIdx.x >= lists[l + 1].blockBegin) l++;
    const SearchList S = lists[l];
    const int b = (int)blockIdx.x - S.blockBegin;
    xg_block B = S.blocks[b];
    const xg_plane &cur = pics[S.curSlot].plane[0];
    int cand[4];
    const int nc = amvp_list(S.mvField, S.mvStride, cur.width, cur.height, B.x, B.y, B.w, B.h, cand);
    uint32_t sad;
    int mv[2];
    const int idx = select_mvp(pics, S.curSlot, S.refSlot, B.x, B.y, B.w, B.h, nc, cand, 0, ctu, false, L, &sad, mv);
    if (threadIdx.x == 0) {
        B.pic = S.curSlot;
        B.ref = S.refSlot;
        B.mvpx = idx >= 0 ? mv[0] : 0;
        B.mvpy = idx >= 0 ? mv[1] : 0;
        S.blocks[b] = B;
    }
}

int32_t amvp_multi_enqueue(hipStream_t st, const xg_picture *dPics, const SearchList *dLists, int nLists, int nBlocks,
                           int ctu)
{
    if (nLists <= 0 || nBlocks <= 0) return XG_OK;
    k_amvp_multi<<<nBlocks, 64, 0, st>>>(dPics, dLists, nLists, ctu);
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
    k_amvp_select<<<n, 64, 0, v.stream>>>(v.dPics, blocks, n, imvIdx, ctuSize, out);
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

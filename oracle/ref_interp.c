/* ref_interp.c -- timing driver for the reference's own interpolation code.
 *
 * Bench and test infrastructure only, like the oracle: nothing in the
 * product links it.  It runs the interpolation share of the 1080p fractional
 * refinement -- for every block and each of its candidates, the luma 8-tap
 * prediction and both chroma 4-tap predictions -- through the reference's
 * 10-bit build of h266_core/source/h266_luma_mc.c and h266_chroma_mc.c,
 * dispatched on (!!fracX, !!fracY) as the encoder's tables dispatch them
 * (h266_encoder/source/h266_inter_prediction.c:143-172 luma, :174-236
 * chroma).  Candidate MVs are clipped to the CU window
 * (h266_encoder/source/h266_enc_init.c:1462-1465) and split into integer and
 * fractional parts (h266_inter_prediction.c:332-342 luma, :664-674 chroma,
 * 1/32 of a chroma sample), as the refinement does.  oracle/Makefile compiles
 * it with the reference sources into oracle/_ref/libxin_ref.so. */
#include <stdint.h>
#include <stddef.h>

#include "xin_typedef.h"
#include "h266_motion_comp.h"

static int32_t clip3i(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

static void luma_pred(const PIXEL *src, intptr_t ss, PIXEL *dst, intptr_t ds, int32_t frac, uint32_t w, uint32_t h)
{
    const int fx = (frac & 15) != 0, fy = (frac >> 4) != 0;
    if (fx && fy) Xin266LumaInterpHorVet(src, ss, dst, ds, frac, w, h, 0);
    else if (fx) Xin266LumaInterpHor(src, ss, dst, ds, frac, w, h, 0);
    else if (fy) Xin266LumaInterpVet(src, ss, dst, ds, frac, w, h, 0);
    else Xin266InterpCopy(src, ss, dst, ds, frac, w, h, 0);
}

static void chroma_pred(const PIXEL *src, intptr_t ss, PIXEL *dst, intptr_t ds, int32_t frac, uint32_t w, uint32_t h)
{
    const int fx = (frac & 31) != 0, fy = (frac >> 5) != 0;
    if (fx && fy) Xin266ChromaInterpHorVet(src, ss, dst, ds, frac, w, h, 0);
    else if (fx) Xin266ChromaInterpHor(src, ss, dst, ds, frac, w, h, 0);
    else if (fy) Xin266ChromaInterpVet(src, ss, dst, ds, frac, w, h, 0);
    else Xin266InterpCopy(src, ss, dst, ds, frac, w, h, 0);
}

/* Blocks [0, nblk) of blk (int32 x, y, w, h, mvx, mvy, mvpx, mvpy each; MVs
 * in 1/16 pel), ncand offsets (1/16 pel) each.  ref planes point at sample
 * (0, 0) and are padded for the CU window.  Each prediction goes to the
 * caller's 64x64 / 32x32 scratch planes (rows of 64 and 32 samples), which
 * hold the last candidate's predictions of the last block afterwards.
 * Returns the block samples covered (sum of w * h), the unit of the metric. */
int64_t ref_RefineInterpShare(const PIXEL *refY, const PIXEL *refU, const PIXEL *refV, intptr_t rsY, intptr_t rsC,
                              int32_t fw, int32_t fh, const int32_t *blk, int32_t nblk, int32_t ncand,
                              const int16_t *offx, const int16_t *offy, int32_t ctuSize, PIXEL *outY, PIXEL *outU,
                              PIXEL *outV)
{
    int64_t px = 0;
    for (int32_t b = 0; b < nblk; b++) {
        const int32_t *B = blk + 8 * (size_t)b;
        const int32_t x = B[0], y = B[1], w = B[2], h = B[3];
        const int32_t mnx = (int16_t)(-x - ctuSize - 8), mny = (int16_t)(-y - ctuSize - 8);
        const int32_t mxx = (int16_t)(fw + (ctuSize - w + 8) - x), mxy = (int16_t)(fh + (ctuSize - h + 8) - y);
        for (int32_t c = 0; c < ncand; c++) {
            const int32_t mvx = clip3i(B[4] + offx[c], mnx * 16, mxx * 16);
            const int32_t mvy = clip3i(B[5] + offy[c], mny * 16, mxy * 16);
            const int32_t rx = (x << 4) + mvx, ry = (y << 4) + mvy;
            luma_pred(refY + (intptr_t)(ry >> 4) * rsY + (rx >> 4), rsY, outY, 64, ((ry & 15) << 4) + (rx & 15),
                      (uint32_t)w, (uint32_t)h);
            const intptr_t co = (intptr_t)(ry >> 5) * rsC + (rx >> 5);
            const int32_t fracC = ((ry & 31) << 5) + (rx & 31);
            chroma_pred(refU + co, rsC, outU, 32, fracC, (uint32_t)w / 2, (uint32_t)h / 2);
            chroma_pred(refV + co, rsC, outV, 32, fracC, (uint32_t)w / 2, (uint32_t)h / 2);
        }
        px += (int64_t)w * h;
    }
    return px;
}

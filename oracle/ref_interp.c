/* ref_interp.c -- timing driver for the reference's own interpolation code.
 *
 * Bench and test infrastructure only, like the oracle: nothing in the
 * product links it.  It runs the interpolation share of the 1080p fractional
 * refinement -- for every block and each of its candidates, the luma 8-tap
 * prediction and both chroma 4-tap predictions -- through the reference's own
 * interpolation functions.  Candidate MVs are clipped to the CU window
 * (h266_encoder/source/h266_enc_init.c:1462-1465) and split into integer and
 * fractional parts (h266_inter_prediction.c:332-342 luma, :664-674 chroma,
 * 1/32 of a chroma sample), as the refinement does.
 *
 * oracle/Makefile compiles it twice with the reference sources:
 *  - oracle/_ref/libxin_ref.so: the 10-bit build (ENABLE_10BIT_ENCODER) of
 *    h266_core/source/h266_luma_mc.c and h266_chroma_mc.c, which has C code
 *    only, dispatched on (!!fracX, !!fracY) as the encoder does
 *    (h266_encoder/source/h266_inter_prediction.c:143-172 luma, :174-236
 *    chroma);
 *  - oracle/_ref/libxin_ref8.so (REF_X86_DISPATCH): the reference's default
 *    build (CMakeLists.txt:33-37 adds _X86_OPT_, nothing adds
 *    ENABLE_10BIT_ENCODER), 8-bit, whose pfXinLumaInterp / pfXinChromaInterp
 *    tables hold SSSE3 and AVX2 code for most block widths. */
#include <stdint.h>
#include <stddef.h>

#include "xin_typedef.h"
#include "h266_motion_comp.h"

static int32_t clip3i(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

#if defined(REF_X86_DISPATCH)
/* The final state of the encoder's interpolation tables, indexed
 * [!!fracX][!!fracY][log2 width] as h266_inter_prediction.c:162 (luma) and
 * :400-410 (chroma, index log2 of the chroma width) index them:
 *  - C entries for every width (h266_encoder/source/h266_func_init.c:424-458
 *    luma, :550-584 chroma);
 *  - replaced under XIN_CPU_SSSE3 (:777) by the SSSE3 entries of :824-849
 *    (luma) and :891-913 (chroma);
 *  - then under XIN_CPU_AVX2 (:1006) by the AVX2 luma entries of :1182-1197.
 * ref_SetDispatch(level) rebuilds the tables with the C entries only (0),
 * with SSSE3 (1) or with SSSE3 and AVX2 (2); levels the host CPU lacks are
 * left out, as the reference's cpuFeature test leaves them out.  Call it
 * before any thread runs ref_RefineInterpShare. */
typedef void (*interp_fn)(const PIXEL *, intptr_t, PIXEL *, intptr_t, SINT32, UINT32, UINT32, UINT32);
static interp_fn luma_tab[2][2][8], chroma_tab[2][2][8];
static int dispatch_level = -1;

static void fill_row(interp_fn t[8], interp_fn f)
{
    for (int i = 0; i < 8; i++) t[i] = f;
}

int32_t ref_SetDispatch(int32_t level)
{
    __builtin_cpu_init();
    if (level >= 1 && !__builtin_cpu_supports("ssse3")) level = 0;
    if (level >= 2 && !__builtin_cpu_supports("avx2")) level = 1;
    fill_row(luma_tab[0][0], Xin266InterpCopy);
    fill_row(luma_tab[1][0], Xin266LumaInterpHor);
    fill_row(luma_tab[0][1], Xin266LumaInterpVet);
    fill_row(luma_tab[1][1], Xin266LumaInterpHorVet);
    fill_row(chroma_tab[0][0], Xin266InterpCopy);
    fill_row(chroma_tab[1][0], Xin266ChromaInterpHor);
    fill_row(chroma_tab[0][1], Xin266ChromaInterpVet);
    fill_row(chroma_tab[1][1], Xin266ChromaInterpHorVet);
    if (level >= 1) {
        interp_fn copy[5] = {Xin266InterpCopy4xH_SSSE3, Xin266InterpCopy8xH_SSSE3, Xin266InterpCopy16xH_SSSE3,
                             Xin266InterpCopy32xH_SSSE3, Xin266InterpCopy64xH_SSSE3};
        for (int i = 0; i < 5; i++) luma_tab[0][0][2 + i] = chroma_tab[0][0][2 + i] = copy[i];
        luma_tab[0][1][2] = Xin266LumaInterpVet4xH_SSSE3;
        luma_tab[0][1][3] = Xin266LumaInterpVet8xH_SSSE3;
        luma_tab[0][1][4] = Xin266LumaInterpVet16xH_SSSE3;
        luma_tab[1][0][2] = Xin266LumaInterpHor4xH_SSSE3;
        luma_tab[1][0][3] = Xin266LumaInterpHor8xH_SSSE3;
        luma_tab[1][0][4] = Xin266LumaInterpHor16xH_SSSE3;
        luma_tab[1][1][2] = Xin266LumaInterpHorVet4xH_SSSE3;
        luma_tab[1][1][3] = Xin266LumaInterpHorVet8xH_SSSE3;
        for (int i = 5; i < 8; i++) {
            luma_tab[0][1][i] = Xin266LumaInterpVetGt16xH_SSSE3;
            luma_tab[1][0][i] = Xin266LumaInterpHorGt16xH_SSSE3;
        }
        for (int i = 4; i < 8; i++) luma_tab[1][1][i] = Xin266LumaInterpHorVetGt8xH_SSSE3;
        chroma_tab[1][0][3] = Xin266ChromaInterpHor8xH_SSSE3;
        chroma_tab[0][1][3] = Xin266ChromaInterpVet8xH_SSSE3;
        chroma_tab[1][1][3] = Xin266ChromaInterpHorVet8xH_SSSE3;
        for (int i = 4; i < 8; i++) {
            chroma_tab[1][0][i] = Xin266ChromaInterpHorGt8xH_SSSE3;
            chroma_tab[0][1][i] = Xin266ChromaInterpVetGt8xH_SSSE3;
            chroma_tab[1][1][i] = Xin266ChromaInterpHorVetGt8xH_SSSE3;
        }
    }
    if (level >= 2) {
        luma_tab[0][0][5] = Xin266InterpCopy32xH_AVX2;
        luma_tab[0][0][6] = Xin266InterpCopy64xH_AVX2;
        luma_tab[0][0][7] = Xin266InterpCopy128xH_AVX2;
        for (int i = 5; i < 8; i++) {
            luma_tab[0][1][i] = Xin266LumaInterpVetGt16xH_AVX2;
            luma_tab[1][0][i] = Xin266LumaInterpHorGt16xH_AVX2;
        }
        for (int i = 4; i < 8; i++) luma_tab[1][1][i] = Xin266LumaInterpHorVetGt8xH_AVX2;
    }
    dispatch_level = level;
    return level;
}

static int log2u(uint32_t w) { return 31 - __builtin_clz(w); }

static void luma_pred(const PIXEL *src, intptr_t ss, PIXEL *dst, intptr_t ds, int32_t frac, uint32_t w, uint32_t h)
{
    luma_tab[(frac & 15) != 0][(frac >> 4) != 0][log2u(w)](src, ss, dst, ds, frac, w, h, 0);
}

static void chroma_pred(const PIXEL *src, intptr_t ss, PIXEL *dst, intptr_t ds, int32_t frac, uint32_t w, uint32_t h)
{
    chroma_tab[(frac & 31) != 0][(frac >> 5) != 0][log2u(w)](src, ss, dst, ds, frac, w, h, 0);
}
#else
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
#endif

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

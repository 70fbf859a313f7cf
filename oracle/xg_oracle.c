/*
 * xg_oracle.c -- CPU oracle (checker) for the VVC block-matching hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product library never uses it.
 *
 * Citations are file:line in the xin26x source tree.  Parity status per part:
 *   interpolation (luma/chroma, S16 variants, bi-avg/weight): pinned against
 *       golden vectors written by the reference C itself (tests/golden/);
 *   SAD / SAD-xN: closed in the reference; closed form, pinned by brute force
 *       and invariants;
 *   SATD / Hadamard 8x8: closed in the reference; HM's published Hadamard
 *       method -- parity unpinned vs the reference;
 *   Exp-Golomb MV bits: closed table; textbook se(v) length -- parity unpinned;
 *   integer search order / ties, 1/16-pel pass: closed or new; build-defined;
 *   forward transforms: closed; H.266 matrices + VTM rounding -- parity
 *       unpinned vs the reference, pinned by closed forms and float DCT/DST.
 */
#include "xg_oracle.h"

#include <stdlib.h>
#include <string.h>

#define BITDEPTH       10
#define PIX_MAX        1023
/* h266_core/include/h266_inter_pred_context.h:31-39 */
#define IF_SHIFT       6
#define IF_PREC        14
#define IF_PREC_OFS    (1 << (IF_PREC - 1))            /* 8192 */
#define IF_HEADROOM    ((IF_PREC - BITDEPTH) > 2 ? (IF_PREC - BITDEPTH) : 2) /* 4 */

static inline int32_t clip_pix(int32_t v) { return v < 0 ? 0 : (v > PIX_MAX ? PIX_MAX : v); }

/* ---------------------------------------------------------------------------
 * Filter tables.  h266_core/source/h266_luma_mc.c:27-70 (VVC luma 1/16-pel
 * 8-tap, 4x4-affine variant, alternative half-pel) and
 * h266_core/source/h266_chroma_mc.c:26-60 (VVC chroma 1/32-pel 4-tap).  These
 * are normative H.266 tables (Tables 27-29), so their values must be equal.
 * ------------------------------------------------------------------------- */
static const int16_t kLuma[16][8] = {
    { 0, 0,   0, 64,  0,   0, 0,  0}, { 0, 1,  -3, 63,  4,  -2, 1,  0},
    {-1, 2,  -5, 62,  8,  -3, 1,  0}, {-1, 3,  -8, 60, 13,  -4, 1,  0},
    {-1, 4, -10, 58, 17,  -5, 1,  0}, {-1, 4, -11, 52, 26,  -8, 3, -1},
    {-1, 3,  -9, 47, 31, -10, 4, -1}, {-1, 4, -11, 45, 34, -10, 4, -1},
    {-1, 4, -11, 40, 40, -11, 4, -1}, {-1, 4, -10, 34, 45, -11, 4, -1},
    {-1, 4, -10, 31, 47,  -9, 3, -1}, {-1, 3,  -8, 26, 52, -11, 4, -1},
    { 0, 1,  -5, 17, 58, -10, 4, -1}, { 0, 1,  -4, 13, 60,  -8, 3, -1},
    { 0, 1,  -3,  8, 62,  -5, 2, -1}, { 0, 1,  -2,  4, 63,  -3, 1,  0}};
static const int16_t kLuma4x4[16][8] = {
    {0, 0,   0, 64,  0,   0, 0, 0}, {0, 1,  -3, 63,  4,  -2, 1, 0},
    {0, 1,  -5, 62,  8,  -3, 1, 0}, {0, 2,  -8, 60, 13,  -4, 1, 0},
    {0, 3, -10, 58, 17,  -5, 1, 0}, {0, 3, -11, 52, 26,  -8, 2, 0},
    {0, 2,  -9, 47, 31, -10, 3, 0}, {0, 3, -11, 45, 34, -10, 3, 0},
    {0, 3, -11, 40, 40, -11, 3, 0}, {0, 3, -10, 34, 45, -11, 3, 0},
    {0, 3, -10, 31, 47,  -9, 2, 0}, {0, 2,  -8, 26, 52, -11, 3, 0},
    {0, 1,  -5, 17, 58, -10, 3, 0}, {0, 1,  -4, 13, 60,  -8, 2, 0},
    {0, 1,  -3,  8, 62,  -5, 1, 0}, {0, 1,  -2,  4, 63,  -3, 1, 0}};
static const int16_t kLumaAltHalf[8] = {0, 3, 9, 20, 20, 9, 3, 0};
static const int16_t kChroma[32][4] = {
    { 0, 64,  0,  0}, {-1, 63,  2,  0}, {-2, 62,  4,  0}, {-2, 60,  7, -1},
    {-2, 58, 10, -2}, {-3, 57, 12, -2}, {-4, 56, 14, -2}, {-4, 55, 15, -2},
    {-4, 54, 16, -2}, {-5, 53, 18, -2}, {-6, 52, 20, -2}, {-6, 49, 24, -3},
    {-6, 46, 28, -4}, {-5, 44, 29, -4}, {-4, 42, 30, -4}, {-4, 39, 33, -4},
    {-4, 36, 36, -4}, {-4, 33, 39, -4}, {-4, 30, 42, -4}, {-4, 29, 44, -5},
    {-4, 28, 46, -6}, {-3, 24, 49, -6}, {-2, 20, 52, -6}, {-2, 18, 53, -5},
    {-2, 16, 54, -4}, {-2, 15, 55, -4}, {-2, 14, 56, -4}, {-2, 12, 57, -3},
    {-2, 10, 58, -2}, {-1,  7, 60, -2}, { 0,  4, 62, -2}, { 0,  2, 63, -1}};

/* Coefficient choice: h266_luma_mc.c:117-128 (same rule in every variant). */
static const int16_t *luma_taps(int32_t phase, uint32_t fi)
{
    if (fi == 1 && phase == 8) return kLumaAltHalf;
    if (fi == 2) return kLuma4x4[phase];
    return kLuma[phase];
}

/* Generic FIR over `taps` samples spaced by `step`, starting at p. */
static inline int32_t fir_u16(const or_pixel *p, intptr_t step, const int16_t *c, int taps)
{
    int32_t s = 0;
    for (int k = 0; k < taps; k++) s += (int32_t)p[k * step] * c[k];
    return s;
}
static inline int32_t fir_s16(const int16_t *p, intptr_t step, const int16_t *c, int taps)
{
    int32_t s = 0;
    for (int k = 0; k < taps; k++) s += (int32_t)p[k * step] * c[k];
    return s;
}

/* ---- PIXEL -> PIXEL single-direction passes ------------------------------
 * h266_luma_mc.c:100-155 (Hor) and :157-214 (Vet):
 *   clip((sum + 32) >> 6).  `step` selects the direction. */
static void pass_u16_to_pix(const or_pixel *src, intptr_t ss, intptr_t step, or_pixel *dst,
                            intptr_t ds, const int16_t *c, int taps, uint32_t w, uint32_t h)
{
    src -= (taps / 2 - 1) * step;
    for (uint32_t y = 0; y < h; y++, src += ss, dst += ds)
        for (uint32_t x = 0; x < w; x++)
            dst[x] = (or_pixel)clip_pix((fir_u16(src + x, step, c, taps) + (1 << (IF_SHIFT - 1))) >> IF_SHIFT);
}

/* ---- PIXEL -> S16 first pass ---------------------------------------------
 * h266_luma_mc.c:216-273 (HorU8S16) and :465-523 (VetU8S16):
 *   (sum - (8192 << hs)) >> hs, hs = 6 - max(2, 14 - bitDepth), stored as int16. */
static void pass_u16_to_s16(const or_pixel *src, intptr_t ss, intptr_t step, int16_t *dst,
                            intptr_t ds, const int16_t *c, int taps, uint32_t w, uint32_t h)
{
    const int32_t hs = IF_SHIFT - IF_HEADROOM;
    const int32_t ofs = -IF_PREC_OFS * (1 << hs);
    src -= (taps / 2 - 1) * step;
    for (uint32_t y = 0; y < h; y++, src += ss, dst += ds)
        for (uint32_t x = 0; x < w; x++)
            dst[x] = (int16_t)((fir_u16(src + x, step, c, taps) + ofs) >> hs);
}

/* ---- S16 -> PIXEL vertical second pass -----------------------------------
 * h266_luma_mc.c:275-336 (VetS16U8):
 *   clip((sum + (1 << (vs-1)) + (8192 << 6)) >> vs), vs = 6 + max(2, 14-bd). */
static void pass_s16_to_pix(const int16_t *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                            const int16_t *c, int taps, uint32_t w, uint32_t h)
{
    const int32_t vs = IF_SHIFT + IF_HEADROOM;
    const int32_t ofs = (1 << (vs - 1)) + (IF_PREC_OFS << IF_SHIFT);
    src -= (taps / 2 - 1) * ss;
    for (uint32_t y = 0; y < h; y++, src += ss, dst += ds)
        for (uint32_t x = 0; x < w; x++)
            dst[x] = (or_pixel)clip_pix((fir_s16(src + x, ss, c, taps) + ofs) >> vs);
}

/* ---- S16 -> S16 vertical second pass -------------------------------------
 * h266_luma_mc.c:338-392 (VetS16S16): sum >> 6, no rounding offset. */
static void pass_s16_to_s16(const int16_t *src, intptr_t ss, int16_t *dst, intptr_t ds,
                            const int16_t *c, int taps, uint32_t w, uint32_t h)
{
    src -= (taps / 2 - 1) * ss;
    for (uint32_t y = 0; y < h; y++, src += ss, dst += ds)
        for (uint32_t x = 0; x < w; x++)
            dst[x] = (int16_t)(fir_s16(src + x, ss, c, taps) >> IF_SHIFT);
}

/* h266_luma_mc.c:72-98 */
void or_InterpCopy(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                   int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)frac; (void)fi;
    for (uint32_t y = 0; y < h; y++, src += ss, dst += ds) memcpy(dst, src, w * sizeof(or_pixel));
}

/* h266_luma_mc.c:100-155: horizontal phase = frac & 15 */
void or_LumaInterpHor(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                      int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    pass_u16_to_pix(src, ss, 1, dst, ds, luma_taps(frac & 15, fi), 8, w, h);
}

/* h266_luma_mc.c:157-214: vertical phase = frac >> 4 */
void or_LumaInterpVet(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                      int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    pass_u16_to_pix(src, ss, ss, dst, ds, luma_taps(frac >> 4, fi), 8, w, h);
}

/* h266_luma_mc.c:216-273 */
void or_LumaInterpHorU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                           int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    pass_u16_to_s16(src, ss, 1, dst, ds, luma_taps(frac & 15, fi), 8, w, h);
}

/* h266_luma_mc.c:465-523 */
void or_LumaInterpVetU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                           int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    pass_u16_to_s16(src, ss, ss, dst, ds, luma_taps(frac >> 4, fi), 8, w, h);
}

/* h266_luma_mc.c:275-336 */
void or_LumaInterpVetS16U8(const int16_t *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                           int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    pass_s16_to_pix(src, ss, dst, ds, luma_taps(frac >> 4, fi), 8, w, h);
}

/* h266_luma_mc.c:338-392 */
void or_LumaInterpVetS16S16(const int16_t *src, intptr_t ss, int16_t *dst, intptr_t ds,
                            int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    pass_s16_to_s16(src, ss, dst, ds, luma_taps(frac >> 4, fi), 8, w, h);
}

/* h266_luma_mc.c:394-428: first pass over h+7 rows from 3 rows above into an
 * int16 buffer of row stride w, then the S16->PIXEL vertical pass. */
void or_LumaInterpHorVet(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                         int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    int16_t *tmp = (int16_t *)malloc(sizeof(int16_t) * w * (h + 7));
    pass_u16_to_s16(src - 3 * ss, ss, 1, tmp, w, luma_taps(frac & 15, fi), 8, w, h + 7);
    pass_s16_to_pix(tmp + 3 * w, w, dst, ds, luma_taps(frac >> 4, fi), 8, w, h);
    free(tmp);
}

/* h266_luma_mc.c:525-559 */
void or_LumaInterpHorVetU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                              int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    int16_t *tmp = (int16_t *)malloc(sizeof(int16_t) * w * (h + 7));
    pass_u16_to_s16(src - 3 * ss, ss, 1, tmp, w, luma_taps(frac & 15, fi), 8, w, h + 7);
    pass_s16_to_s16(tmp + 3 * w, w, dst, ds, luma_taps(frac >> 4, fi), 8, w, h);
    free(tmp);
}

/* h266_luma_mc.c:430-463: (s << 4) - 8192 as int16 */
void or_InterpCopyU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                        int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)frac; (void)fi;
    for (uint32_t y = 0; y < h; y++, src += ss, dst += ds)
        for (uint32_t x = 0; x < w; x++)
            dst[x] = (int16_t)((int16_t)(src[x] << IF_HEADROOM) - IF_PREC_OFS);
}

/* h266_luma_mc.c:561-593: shift = max(2, 14-bd) + 1, offset = (1<<(shift-1)) + 2*8192 */
void or_InterpAvgS16U8(const int16_t *s0, intptr_t st0, const int16_t *s1, intptr_t st1,
                       or_pixel *dst, intptr_t ds, uint32_t w, uint32_t h)
{
    const int32_t sh = IF_HEADROOM + 1;
    const int32_t ofs = (1 << (sh - 1)) + 2 * IF_PREC_OFS;
    for (uint32_t y = 0; y < h; y++, s0 += st0, s1 += st1, dst += ds)
        for (uint32_t x = 0; x < w; x++)
            dst[x] = (or_pixel)clip_pix(((int32_t)s0[x] + s1[x] + ofs) >> sh);
}

/* h266_luma_mc.c:595-629 with XIN_BCW_LOG_WGT_BASE = 3 (h266_core/include/h266_constant.h:24) */
void or_InterpWeightS16U8(const int16_t *s0, intptr_t st0, const int16_t *s1, intptr_t st1,
                          or_pixel *dst, intptr_t ds, uint32_t w, uint32_t h,
                          int32_t wA, int32_t wB)
{
    const int32_t sh = IF_HEADROOM + 3;
    const int32_t ofs = (1 << (sh - 1)) + (IF_PREC_OFS << 3);
    for (uint32_t y = 0; y < h; y++, s0 += st0, s1 += st1, dst += ds)
        for (uint32_t x = 0; x < w; x++)
            dst[x] = (or_pixel)clip_pix(((int32_t)s0[x] * wA + (int32_t)s1[x] * wB + ofs) >> sh);
}

/* ---- chroma: h266_chroma_mc.c, same structure with 4 taps and 32 phases ---- */
/* :62-100 */
void or_ChromaInterpHor(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                        int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    pass_u16_to_pix(src, ss, 1, dst, ds, kChroma[frac & 31], 4, w, h);
}
/* :102-140 */
void or_ChromaInterpVet(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                        int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    pass_u16_to_pix(src, ss, ss, dst, ds, kChroma[frac >> 5], 4, w, h);
}
/* :142-183 */
void or_ChromaInterpHorU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                             int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    pass_u16_to_s16(src, ss, 1, dst, ds, kChroma[frac & 31], 4, w, h);
}
/* :305-346 */
void or_ChromaInterpVetU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                             int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    pass_u16_to_s16(src, ss, ss, dst, ds, kChroma[frac >> 5], 4, w, h);
}
/* :185-228 */
void or_ChromaInterpVetS16U8(const int16_t *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                             int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    pass_s16_to_pix(src, ss, dst, ds, kChroma[frac >> 5], 4, w, h);
}
/* :230-267 */
void or_ChromaInterpVetS16S16(const int16_t *src, intptr_t ss, int16_t *dst, intptr_t ds,
                              int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    pass_s16_to_s16(src, ss, dst, ds, kChroma[frac >> 5], 4, w, h);
}
/* :269-303: first pass over h+3 rows from 1 row above */
void or_ChromaInterpHorVet(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                           int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    int16_t *tmp = (int16_t *)malloc(sizeof(int16_t) * w * (h + 3));
    pass_u16_to_s16(src - ss, ss, 1, tmp, w, kChroma[frac & 31], 4, w, h + 3);
    pass_s16_to_pix(tmp + w, w, dst, ds, kChroma[frac >> 5], 4, w, h);
    free(tmp);
}
/* :348-384 */
void or_ChromaInterpHorVetU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                                int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    int16_t *tmp = (int16_t *)malloc(sizeof(int16_t) * w * (h + 3));
    pass_u16_to_s16(src - ss, ss, 1, tmp, w, kChroma[frac & 31], 4, w, h + 3);
    pass_s16_to_s16(tmp + w, w, dst, ds, kChroma[frac >> 5], 4, w, h);
    free(tmp);
}

/* ---- dispatch as the encoder does it ---- */
/* h266_inter_prediction.c:143-172 with the table of h266_func_init.c:424-458:
 * [0][0] Copy, [1][0] Hor, [0][1] Vet, [1][1] HorVet. */
void or_LumaInterpolation(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                          int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    const int fx = (frac & 15) != 0, fy = ((frac >> 4) & 15) != 0;
    if (fx && fy) or_LumaInterpHorVet(src, ss, dst, ds, frac, w, h, fi);
    else if (fx)  or_LumaInterpHor(src, ss, dst, ds, frac, w, h, fi);
    else if (fy)  or_LumaInterpVet(src, ss, dst, ds, frac, w, h, fi);
    else          or_InterpCopy(src, ss, dst, ds, frac, w, h, fi);
}

/* h266_inter_prediction.c:78-141 */
void or_LumaInterpolationS16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                             int32_t frac, uint32_t w, uint32_t h, uint32_t fi)
{
    if (!frac)              or_InterpCopyU8S16(src, ss, dst, ds, frac, w, h, fi);
    else if (!(frac & 15))  or_LumaInterpVetU8S16(src, ss, dst, ds, frac, w, h, fi);
    else if (!(frac & 240)) or_LumaInterpHorU8S16(src, ss, dst, ds, frac, w, h, fi);
    else                    or_LumaInterpHorVetU8S16(src, ss, dst, ds, frac, w, h, fi);
}

/* h266_inter_prediction.c:174-236 */
void or_ChromaInterpolation(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                            int32_t frac, uint32_t w, uint32_t h)
{
    if (!frac)                  or_InterpCopy(src, ss, dst, ds, frac, w, h, 0);
    else if (!(frac & (31 << 5))) or_ChromaInterpHor(src, ss, dst, ds, frac, w, h, 0);
    else if (!(frac & 31))      or_ChromaInterpVet(src, ss, dst, ds, frac, w, h, 0);
    else                        or_ChromaInterpHorVet(src, ss, dst, ds, frac, w, h, 0);
}

/* h266_inter_prediction.c:238-300 */
void or_ChromaInterpolationS16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                               int32_t frac, uint32_t w, uint32_t h)
{
    if (!frac)                  or_InterpCopyU8S16(src, ss, dst, ds, frac, w, h, 0);
    else if (!(frac & (31 << 5))) or_ChromaInterpHorU8S16(src, ss, dst, ds, frac, w, h, 0);
    else if (!(frac & 31))      or_ChromaInterpVetU8S16(src, ss, dst, ds, frac, w, h, 0);
    else                        or_ChromaInterpHorVetU8S16(src, ss, dst, ds, frac, w, h, 0);
}

/* ---------------------------------------------------------------------------
 * Distortion.  The reference binds these to closed functions
 * (h26x_core/include/h26x_compute_dist.h:27-150, 400-735, 806-937); the
 * signature follows XinComputeDist (h266_encoder/include/h266_func_struct.h:56-63).
 * ------------------------------------------------------------------------- */
uint32_t or_ComputeSad(const or_pixel *in, intptr_t is, const or_pixel *pr, intptr_t ps,
                       uint32_t w, uint32_t h)
{
    uint32_t s = 0;
    for (uint32_t y = 0; y < h; y++, in += is, pr += ps)
        for (uint32_t x = 0; x < w; x++) s += (uint32_t)abs((int)in[x] - (int)pr[x]);
    return s;
}

/* XinComputeSadx8 (h266_func_struct.h:87-94): one input against 8 positions. */
void or_ComputeSadx8(const or_pixel *in, intptr_t is, const or_pixel *const ref[8], intptr_t rs,
                     uint32_t w, uint32_t h, uint32_t sad[8])
{
    for (int k = 0; k < 8; k++) sad[k] = or_ComputeSad(in, is, ref[k], rs, w, h);
}

/* XinComputeSadx5 (h266_func_struct.h:106-117) */
void or_ComputeSadx5(const or_pixel *in, intptr_t is, const or_pixel *r0, const or_pixel *r1,
                     const or_pixel *r2, const or_pixel *r3, const or_pixel *r4, intptr_t rs,
                     uint32_t w, uint32_t h, uint32_t sad[5])
{
    const or_pixel *r[5] = {r0, r1, r2, r3, r4};
    for (int k = 0; k < 5; k++) sad[k] = or_ComputeSad(in, is, r[k], rs, w, h);
}

/* XinComputeSadx3 (h266_func_struct.h:119-128) */
void or_ComputeSadx3(const or_pixel *in, intptr_t is, const or_pixel *r0, const or_pixel *r1,
                     const or_pixel *r2, intptr_t rs, uint32_t w, uint32_t h, uint32_t sad[3])
{
    const or_pixel *r[3] = {r0, r1, r2};
    for (int k = 0; k < 3; k++) sad[k] = or_ComputeSad(in, is, r[k], rs, w, h);
}

/* Bi-prediction distortions XinComputeAvgDist / XinComputeAvgSadx8
 * (h266_encoder/include/h266_func_struct.h:65-73, 96-104), bound to the closed
 * Xin26xComputeAvgSad / AvgSatd / AvgSadx8 (h26x_core/include/h26x_compute_dist.h:
 * 153-243, 735-795, 824-900) and called on the two uni-directional predictions
 * of a B-frame PU (h266_encoder/source/h266_analyze_inter_cu.c:720, 817, 1214).
 * The distortion is that of the input against the rounded average
 * (a + b + 1) >> 1 of the two predictions.  The averaging rounding is
 * parity unpinned vs the closed library; with a == b each reduces to
 * the plain SAD / SATD exactly. */
static void avg_block(const or_pixel *a, const or_pixel *b, intptr_t rs, uint32_t w, uint32_t h, or_pixel *out)
{
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++)
            out[y * w + x] = (or_pixel)(((uint32_t)a[y * rs + x] + b[y * rs + x] + 1) >> 1);
}

uint32_t or_ComputeAvgSad(const or_pixel *in, intptr_t is, const or_pixel *a, const or_pixel *b, intptr_t rs,
                          uint32_t w, uint32_t h)
{
    uint32_t s = 0;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++)
            s += (uint32_t)abs((int)in[y * is + x] - (int)(((uint32_t)a[y * rs + x] + b[y * rs + x] + 1) >> 1));
    return s;
}

void or_ComputeAvgSadx8(const or_pixel *in, intptr_t is, const or_pixel *const a[8], const or_pixel *const b[8],
                        intptr_t rs, uint32_t w, uint32_t h, uint32_t sad[8])
{
    for (int k = 0; k < 8; k++) sad[k] = or_ComputeAvgSad(in, is, a[k], b[k], rs, w, h);
}

/* In-place 1-D Walsh-Hadamard transform of n (power of two) values spaced by step. */
static void wht_1d(int32_t *v, int n, int step)
{
    for (int len = 1; len < n; len <<= 1)
        for (int i = 0; i < n; i += len << 1)
            for (int j = i; j < i + len; j++) {
                int32_t a = v[j * step], b = v[(j + len) * step];
                v[j * step] = a + b;
                v[(j + len) * step] = a - b;
            }
}

/* Sum of |2-D WHT coefficients| of an n x n difference tile. */
static uint32_t hadamard_tile(const or_pixel *in, intptr_t is, const or_pixel *pr, intptr_t ps, int n)
{
    int32_t d[64];
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) d[y * n + x] = (int32_t)in[y * is + x] - (int32_t)pr[y * ps + x];
    for (int y = 0; y < n; y++) wht_1d(d + y * n, n, 1);
    for (int x = 0; x < n; x++) wht_1d(d + x, n, n);
    uint32_t s = 0;
    for (int i = 0; i < n * n; i++) s += (uint32_t)abs(d[i]);
    return s;
}

/* SATD: HM's published method (TComRdCost::xCalcHADs8x8/4x4/2x2 and getHADs):
 * 8x8 tiles when both sides are multiples of 8, else 4x4, else 2x2; per-tile
 * normalisation (s+2)>>2 (8x8), (s+1)>>1 (4x4), none (2x2).  The coefficient
 * order of the transform does not matter because only |coef| is summed.
 * Parity unpinned vs the reference's closed Xin26xComputeSatd. */
uint32_t or_ComputeSatd(const or_pixel *in, intptr_t is, const or_pixel *pr, intptr_t ps,
                        uint32_t w, uint32_t h)
{
    uint32_t s = 0;
    if (w % 8 == 0 && h % 8 == 0) {
        for (uint32_t y = 0; y < h; y += 8)
            for (uint32_t x = 0; x < w; x += 8)
                s += (hadamard_tile(in + y * is + x, is, pr + y * ps + x, ps, 8) + 2) >> 2;
    } else if (w % 4 == 0 && h % 4 == 0) {
        for (uint32_t y = 0; y < h; y += 4)
            for (uint32_t x = 0; x < w; x += 4)
                s += (hadamard_tile(in + y * is + x, is, pr + y * ps + x, ps, 4) + 1) >> 1;
    } else {
        for (uint32_t y = 0; y < h; y += 2)
            for (uint32_t x = 0; x < w; x += 2)
                s += hadamard_tile(in + y * is + x, is, pr + y * ps + x, ps, 2);
    }
    return s;
}

/* XinComputeHada8x8 (h266_func_struct.h:448-451): Hadamard magnitude of one
 * 8x8 block of samples, defined here as the 8x8 SATD against a zero block.
 * Parity unpinned vs the closed Xin26xComputeHada8x8. */
uint32_t or_ComputeAvgSatd(const or_pixel *in, intptr_t is, const or_pixel *a, const or_pixel *b, intptr_t rs,
                           uint32_t w, uint32_t h)
{
    or_pixel avg[128 * 128];
    if (w > 128 || h > 128) return UINT32_MAX;
    avg_block(a, b, rs, w, h, avg);
    return or_ComputeSatd(in, is, avg, w, w, h);
}

uint32_t or_ComputeHada8x8(const or_pixel *in, intptr_t is)
{
    static const or_pixel zero[8] = {0};
    return (hadamard_tile(in, is, zero, 0, 8) + 2) >> 2;
}

/* ---- rate terms ----------------------------------------------------------
 * CALC_MVD_BIT (common/include/h26x_definition.h:176) indexes the closed
 * ExpGolombBits table.  Textbook signed Exp-Golomb length se(v):
 * k = v > 0 ? 2v-1 : -2v, length = 2*floor(log2(k+1)) + 1.  Parity unpinned. */
uint32_t or_ExpGolombBits(int32_t v)
{
    uint32_t k = v > 0 ? (uint32_t)(2 * v - 1) : (uint32_t)(-2 * v);
    uint32_t n = k + 1, lg = 0;
    while (n >>= 1) lg++;
    return 2 * lg + 1;
}

uint32_t or_MvdBits(int32_t dx, int32_t dy) { return or_ExpGolombBits(dx) + or_ExpGolombBits(dy); }

/* sadLambda table (h266_encoder/source/h266_enc_init.c:86-105) indexed by
 * qp + 12*is10Bit (:384).  Every entry is round(64 * 2^(i/6)). */
uint32_t or_SadLambda(int32_t qp, int32_t is10bit)
{
    static const uint32_t tab[64] = {
        64, 72, 81, 91, 102, 114, 128, 144, 161, 181, 203, 228, 256, 287, 323, 362,
        406, 456, 512, 575, 645, 724, 813, 912, 1024, 1149, 1290, 1448, 1625, 1825, 2048, 2299,
        2580, 2896, 3251, 3649, 4096, 4598, 5161, 5793, 6502, 7298, 8192, 9195, 10321, 11585, 13004, 14596,
        16384, 18390, 20643, 23170, 26008, 29193, 32768, 36781, 41285, 46341, 52016, 58386, 65536, 73562, 82570, 92682};
    int32_t i = qp + 12 * (is10bit ? 1 : 0);
    if (i < 0) i = 0;
    if (i > 63) i = 63;
    return tab[i];
}

/* h266_get_neighbour_mv.c:245-290 (shift < 0 branch): rounding right shift
 * v >= 0 ? (v + ofs - 1) >> s : (v + ofs) >> s, ofs = 1 << (s-1). */
int32_t or_MvRoundShift(int32_t v, int32_t s)
{
    if (s <= 0) return v * (1 << (-s));
    int32_t ofs = 1 << (s - 1);
    return v >= 0 ? (v + ofs - 1) >> s : (v + ofs) >> s;
}

/* ---- motion search -------------------------------------------------------- */
static inline uint32_t me_cost(uint32_t sad, int32_t mvq_x, int32_t mvq_y, const or_me *me)
{
    /* cost = (SAD << XIN_COST_FRACTION) + bits(mvd) * sadLambda,
     * h266_sub_pel_refine.c:155-156, XIN_COST_FRACTION = 8 (h26x_definition.h:88) */
    return (sad << 8) + or_MvdBits(mvq_x - me->predMvX, mvq_y - me->predMvY) * me->sadLambda;
}

/* Full search (closed Xin26xMotionEstFull, h26x_core/include/h26x_motion_est.h:26).
 * Build-defined: window centred on the predictor rounded to integer pel the way
 * Xin266FillMeCand rounds (h266_analyze_inter_cu.c:271-272), +/- searchWidth x
 * searchHeight intersected with [minMv, maxMv]; raster scan, dy outer and dx
 * inner ascending; strict < keeps the earliest.  Parity unpinned. */
void or_MotionEstFull(or_me *me)
{
    const int32_t cx = (me->predMvX + 2) >> 2, cy = (me->predMvY + 2) >> 2;
    int32_t x0 = cx - me->searchWidth, x1 = cx + me->searchWidth;
    int32_t y0 = cy - me->searchHeight, y1 = cy + me->searchHeight;
    if (x0 < me->minMvX) x0 = me->minMvX;
    if (x1 > me->maxMvX) x1 = me->maxMvX;
    if (y0 < me->minMvY) y0 = me->minMvY;
    if (y1 > me->maxMvY) y1 = me->maxMvY;
    uint32_t best = UINT32_MAX, bestSad = UINT32_MAX;
    int32_t bx = me->bestMvX >> 2, by = me->bestMvY >> 2;
    for (int32_t my = y0; my <= y1; my++)
        for (int32_t mx = x0; mx <= x1; mx++) {
            uint32_t sad = or_ComputeSad(me->input, me->inputStride,
                                         me->ref + my * me->refStride + mx, me->refStride,
                                         me->width, me->height);
            uint32_t c = me_cost(sad, mx * 4, my * 4, me);
            if (c < best) { best = c; bestSad = sad; bx = mx; by = my; }
        }
    me->bestMvX = (int16_t)(bx * 4);
    me->bestMvY = (int16_t)(by * 4);
    me->bestCost = best;
    me->bestSad = bestSad;
}

/* Diamond search (closed Xin26xMotionEstDia, h26x_motion_est.h:35).
 * Build-defined: evaluate meCand[0..candNum) in list order (each clipped to
 * [minMv, maxMv]), then repeat a small-diamond step around the best point with
 * the neighbour order L, R, T, B of mvOffset (h266_sub_pel_refine.c:57-60),
 * keeping points inside [minMv, maxMv] and inside the search window around the
 * rounded predictor, until no neighbour is strictly better or
 * searchWidth + searchHeight steps were taken.  Parity unpinned. */
void or_MotionEstDia(or_me *me)
{
    static const int8_t dxs[4] = {-1, 1, 0, 0}, dys[4] = {0, 0, -1, 1};
    const int32_t cx = (me->predMvX + 2) >> 2, cy = (me->predMvY + 2) >> 2;
    uint32_t best = UINT32_MAX, bestSad = UINT32_MAX;
    int32_t bx = 0, by = 0;
    for (uint32_t i = 0; i < me->candNum; i++) {
        int32_t mx = me->candX[i], my = me->candY[i];
        mx = mx < me->minMvX ? me->minMvX : (mx > me->maxMvX ? me->maxMvX : mx);
        my = my < me->minMvY ? me->minMvY : (my > me->maxMvY ? me->maxMvY : my);
        uint32_t sad = or_ComputeSad(me->input, me->inputStride,
                                     me->ref + my * me->refStride + mx, me->refStride,
                                     me->width, me->height);
        uint32_t c = me_cost(sad, mx * 4, my * 4, me);
        if (c < best) { best = c; bestSad = sad; bx = mx; by = my; }
    }
    const int32_t steps = me->searchWidth + me->searchHeight;
    for (int32_t it = 0; it < steps; it++) {
        int32_t nx = bx, ny = by;
        uint32_t nb = best, nsad = bestSad;
        for (int k = 0; k < 4; k++) {
            int32_t mx = bx + dxs[k], my = by + dys[k];
            if (mx < me->minMvX || mx > me->maxMvX || my < me->minMvY || my > me->maxMvY) continue;
            if (mx < cx - me->searchWidth || mx > cx + me->searchWidth ||
                my < cy - me->searchHeight || my > cy + me->searchHeight) continue;
            uint32_t sad = or_ComputeSad(me->input, me->inputStride,
                                         me->ref + my * me->refStride + mx, me->refStride,
                                         me->width, me->height);
            uint32_t c = me_cost(sad, mx * 4, my * 4, me);
            if (c < nb) { nb = c; nsad = sad; nx = mx; ny = my; }
        }
        if (nx == bx && ny == by) break;
        bx = nx; by = ny; best = nb; bestSad = nsad;
    }
    me->bestMvX = (int16_t)(bx * 4);
    me->bestMvY = (int16_t)(by * 4);
    me->bestCost = best;
    me->bestSad = bestSad;
}

/* h266_sub_pel_refine.c:47-55 neighbour order L, R, T, B, TL, TR, BL, BR. */
static const int8_t kRingX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
static const int8_t kRingY[8] = {0, 0, -1, 1, -1, -1, 1, 1};

/* SAD of the block predicted at a 1/16-pel MV relative to me->ref, through the
 * luma interpolation table (h266_sub_pel_refine.c:134-153). */
static uint32_t sad_at_mv16(const or_me *me, int32_t mv16x, int32_t mv16y, or_pixel *buf)
{
    int32_t frac = ((mv16y & 15) << 4) + (mv16x & 15);
    or_LumaInterpolation(me->ref + me->refStride * (mv16y >> 4) + (mv16x >> 4), me->refStride,
                         buf, 128, frac, me->width, me->height, 0);
    return or_ComputeSad(me->input, me->inputStride, buf, 128, me->width, me->height);
}

/* h266_sub_pel_refine.c:81-212 (Xin266FullSubPelMe): half-pel ring (+/-2
 * quarter-pel) then quarter-pel ring (+/-1) around the integer best, strict <.
 * filterIndex is XIN_INTERP_DEF_FILTER (0): the reference call site passes 8
 * arguments to a 9-parameter function, leaving filterIndex undefined. */
void or_FullSubPelMe(or_me *me)
{
    or_pixel *buf = (or_pixel *)malloc(sizeof(or_pixel) * 128 * 128);
    int32_t mx = me->bestMvX, my = me->bestMvY;
    uint32_t best = me->bestCost;
    if ((mx >> 2) < me->minMvX || (my >> 2) < me->minMvY || (mx >> 2) > me->maxMvX ||
        (my >> 2) > me->maxMvY) {
        free(buf);
        return;
    }
    for (int pass = 0; pass < 2; pass++) {
        const int32_t step = pass == 0 ? 2 : 1;
        const int32_t cx = mx, cy = my;
        for (int k = 0; k < 8; k++) {
            int32_t tx = cx + kRingX[k] * step, ty = cy + kRingY[k] * step;
            uint32_t sad = sad_at_mv16(me, tx * 4, ty * 4, buf);
            uint32_t c = me_cost(sad, tx, ty, me);
            if (c < best) { best = c; mx = tx; my = ty; }
        }
    }
    me->bestMvX = (int16_t)mx;
    me->bestMvY = (int16_t)my;
    me->bestCost = best;
    free(buf);
}

/* h266_analyze_inter_cu.c:133-293 (Xin266FillMeCand): meCand[0] = (0,0); the
 * predictor and the extra quarter-pel candidates are rounded to integer pel by
 * (v + 2) >> 2 (:271-272) and appended unless already present (:265-291). */
void or_FillMeCand(int16_t predX, int16_t predY, const int16_t *qx, const int16_t *qy,
                   uint32_t n, int16_t *outX, int16_t *outY, uint32_t *outNum)
{
    uint32_t cnt = 1;
    outX[0] = 0;
    outY[0] = 0;
    for (uint32_t i = 0; i <= n && i < 11; i++) {
        int32_t vx = i == 0 ? predX : qx[i - 1];
        int32_t vy = i == 0 ? predY : qy[i - 1];
        int16_t cx = (int16_t)((vx + 2) >> 2), cy = (int16_t)((vy + 2) >> 2);
        int dup = 0;
        for (uint32_t j = 0; j < cnt; j++)
            if (outX[j] == cx && outY[j] == cy) { dup = 1; break; }
        if (!dup) { outX[cnt] = cx; outY[cnt] = cy; cnt++; }
    }
    *outNum = cnt;
}

/* ---------------------------------------------------------------------------
 * Forward transforms.  The reference's matrices and kernels are closed
 * (h26x_core/include/h26x_trans_context.h:53-67, h26x_forward_1d_trans.h).
 * Matrices: ITU-T H.266 clause 8.7.4.  DCT-II of every size takes rows of the
 * 64-point matrix (T_N[k][n] = T_64[k * 64/N][n]); the 64-point matrix is built
 * from the cosine-magnitude tables of each dyadic level with the DCT-II sign
 * pattern.  DST-VII from its sine-magnitude table; DCT-VIII as the
 * row-alternating, column-reversed DST-VII.  Parity unpinned vs the reference.
 * ------------------------------------------------------------------------- */
static const int32_t kCos64[32] = {91, 90, 90, 90, 88, 87, 86, 84, 83, 81, 79, 77, 73, 71, 69, 65,
                                   62, 59, 56, 52, 48, 44, 41, 37, 33, 28, 24, 20, 15, 11, 7, 2};
static const int32_t kCos32[16] = {90, 90, 88, 85, 82, 78, 73, 67, 61, 54, 46, 38, 31, 22, 13, 4};
static const int32_t kCos16[8] = {90, 87, 80, 70, 57, 43, 25, 9};
static const int32_t kCos8[4] = {89, 75, 50, 18};
static const int32_t kCos4[2] = {83, 36};
static const int32_t kCos2[1] = {64};
static const int32_t kSin4[4] = {29, 55, 74, 84};
static const int32_t kSin8[8] = {17, 32, 46, 60, 71, 78, 85, 86};
static const int32_t kSin16[16] = {8, 17, 25, 33, 40, 48, 55, 62, 68, 73, 77, 81, 85, 87, 88, 88};
static const int32_t kSin32[32] = {4, 9, 13, 17, 21, 26, 30, 34, 38, 42, 46, 50, 53, 56, 60, 63,
                                   66, 68, 72, 74, 77, 78, 80, 82, 84, 85, 86, 87, 88, 89, 90, 90};

/* Entry (k, n) of the 64-point DCT-II: 64 for k = 0; otherwise write
 * k = 2^e * k' (k' odd).  The level size L = 64 / 2^e holds magnitudes of
 * cos((2i+1) pi / (2L)), i < L/2; the angle index m = (2n+1) k' mod 4L picks
 * the magnitude and the quadrant picks the sign. */
static int32_t dct2_64(int k, int n)
{
    if (k == 0) return 64;
    int kk = k, L = 64;
    while (!(kk & 1)) { kk >>= 1; L >>= 1; }
    const int32_t *mag = L == 64 ? kCos64 : L == 32 ? kCos32 : L == 16 ? kCos16
                       : L == 8 ? kCos8 : L == 4 ? kCos4 : kCos2;
    int m = ((2 * n + 1) * kk) % (4 * L);
    int q = m / L;
    switch (q) {
    case 0: return mag[(m - 1) / 2];
    case 1: return -mag[(2 * L - m - 1) / 2];
    case 2: return -mag[(m - 2 * L - 1) / 2];
    default: return mag[(4 * L - m - 1) / 2];
    }
}

/* DST-VII (k, n) of an N-point transform: sign and magnitude of
 * sin(pi (2k+1)(n+1) / (2N+1)); mag[j-1] holds |sin(j pi / (2N+1))| scaled. */
static int32_t dst7(int N, int k, int n)
{
    const int32_t *mag = N == 4 ? kSin4 : N == 8 ? kSin8 : N == 16 ? kSin16 : kSin32;
    const int P = 2 * N + 1;
    int m = ((2 * k + 1) * (n + 1)) % (2 * P);
    int sgn = 1;
    if (m == 0 || m == P) return 0;
    if (m > P) { m -= P; sgn = -1; }
    int j = m <= N ? m : P - m;
    return sgn * mag[j - 1];
}

int or_TrMatrix(int type, int n, int32_t *out)
{
    if (type == OR_DCT2) {
        if (n < 2 || n > 64 || (n & (n - 1))) return -1;
        for (int k = 0; k < n; k++)
            for (int i = 0; i < n; i++) out[k * n + i] = dct2_64(k * (64 / n), i);
        return 0;
    }
    if (n != 4 && n != 8 && n != 16 && n != 32) return -1;
    for (int k = 0; k < n; k++)
        for (int i = 0; i < n; i++)
            out[k * n + i] = type == OR_DST7 ? dst7(n, k, i)
                           : ((k & 1) ? -1 : 1) * dst7(n, k, n - 1 - i);
    return 0;
}

/* Number of leading outputs a 1-D forward pass keeps (VVC zero-out; the
 * quantiser reads only min(W,32) x min(H,32), h266_trans_quant.c:208-223). */
static int kept_outputs(int type, int n)
{
    if (type == OR_DCT2) return n == 64 ? 32 : n;
    return n == 32 ? 16 : n;
}

/* 1-D forward pass with the XinFor1dDct contract (h266_func_struct.h:198-204,
 * layout from the two calls at h266_forward_trans.c:114-128):
 *   out[k*os + j] = (sum_i T[k][i] * in[j*is + i] + 2^(shift-1)) >> shift
 * for j < line, k < kept; zero for k >= kept. */
void or_FDct1d(int type, int n, const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
               int32_t line, int32_t shift)
{
    int32_t T[64 * 64];
    if (or_TrMatrix(type, n, T)) return;
    const int keep = kept_outputs(type, n);
    const int64_t rnd = shift > 0 ? ((int64_t)1 << (shift - 1)) : 0;
    for (int32_t j = 0; j < line; j++)
        for (int k = 0; k < n; k++) {
            if (k >= keep) { out[k * os + j] = 0; continue; }
            int64_t s = 0;
            for (int i = 0; i < n; i++) s += (int64_t)T[k * n + i] * in[j * is + i];
            out[k * os + j] = (or_coeff)((s + rnd) >> shift);
        }
}

/* XinFor2dDct2 (h266_func_struct.h:206-215): two DCT-II passes through tmp. */
void or_FDct2WxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
                 uint32_t w, uint32_t h, or_coeff *tmp, int32_t s1, int32_t s2)
{
    or_FDct1d(OR_DCT2, (int)w, in, is, tmp, h, (int32_t)h, s1);
    or_FDct1d(OR_DCT2, (int)h, tmp, h, out, os, (int32_t)w, s2);
}

/* Residual as COEFF, XinBlockSubForDct (h26x_core/include/h26x_block_utility.h:151-159),
 * called at h266_sub_dct.c:74-82. */
void or_BlockSubForDct(const or_pixel *in, intptr_t is, const or_pixel *pred, intptr_t ps,
                       or_coeff *diff, intptr_t dstride, uint32_t w, uint32_t h)
{
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++)
            diff[y * dstride + x] = (or_coeff)in[y * is + x] - (or_coeff)pred[y * ps + x];
}

/* mtsIdx2txType (closed, h26x_trans_context.h:66) as the H.266 mts_idx
 * semantics fix it: [hor, ver] for DCT2_DCT2, SKIP, DST7_DST7, DCT8_DST7,
 * DST7_DCT8, DCT8_DCT8 (enum order h26x_trans_context.h:42-51). */
void or_MtsTypes(uint32_t mtsIdx, int32_t *horTx, int32_t *verTx)
{
    static const int8_t tab[6][2] = {{OR_DCT2, OR_DCT2}, {OR_SKIP, OR_SKIP}, {OR_DST7, OR_DST7},
                                     {OR_DCT8, OR_DST7}, {OR_DST7, OR_DCT8}, {OR_DCT8, OR_DCT8}};
    *horTx = tab[mtsIdx < 6 ? mtsIdx : 0][0];
    *verTx = tab[mtsIdx < 6 ? mtsIdx : 0][1];
}

/* h266_forward_trans.c:45-152 (Xin266FDctWxH) without the transform-skip branch. */
void or_FDctWxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
                uint32_t w, uint32_t h, uint32_t mtsIdx, int isIntraLuma, int enableMts, int bitDepth)
{
    int32_t hor, ver;
    int lgW = 0, lgH = 0;
    while ((1u << lgW) < w) lgW++;
    while ((1u << lgH) < h) lgH++;
    or_MtsTypes(mtsIdx, &hor, &ver);
    const int32_t s1 = lgW + bitDepth + 6 - 15, s2 = lgH + 6;   /* :72-73 */
    if (isIntraLuma && enableMts && mtsIdx == 0) {               /* :75-86 */
        if (w >= 4 && w <= 16) hor = OR_DST7;
        if (h >= 4 && h <= 16) ver = OR_DST7;
    }
    if (hor == OR_SKIP) return;
    or_coeff *tmp = (or_coeff *)malloc(sizeof(or_coeff) * w * h);
    if (lgW > 0 && lgH > 0) {                                    /* :88-130 */
        or_FDct1d(hor, (int)w, in, is, tmp, h, (int32_t)h, s1);
        or_FDct1d(ver, (int)h, tmp, h, out, os, (int32_t)w, s2);
    } else if (lgW > 0) {
        or_FDct1d(hor, (int)w, in, is, out, os, (int32_t)h, s1);
    } else {
        or_FDct1d(ver, (int)h, in, is, out, os, (int32_t)w, s2);
    }
    free(tmp);
}

/* ---------------------------------------------------------------------------
 * Inverse transforms.  The reference's 1-D kernels are closed
 * (h26x_core/include/h26x_inverse_1d_trans.h:23-200); the 2-D dispatch,
 * transform skip and reconstruction call order are source
 * (h266_encoder/source/h266_inverse_trans.c:45-229, h266_idct_add.c:45-104).
 * 1-D kernel restated from ITU-T H.266 clause 8.7.4.5 (transformation
 * process): y[i] = sum_{k < nonZeroS} T[k][i] x[k], with nonZeroS the
 * zero-out bound of clause 8.7.4.2 (32 for DCT-II-64, 16 for 32-point
 * DST-VII/DCT-VIII); each pass rounds, shifts and clips to
 * [-32768, 32767] (clause 8.7.4.2 clips the first pass to coeffMin/coeffMax;
 * the output is COEFF-typed and saturates at the same bounds).
 * ------------------------------------------------------------------------- */
static int32_t clip16(int64_t v) { return (int32_t)(v < -32768 ? -32768 : v > 32767 ? 32767 : v); }

/* XinInv1dDct (h266_func_struct.h:280-287), layout from the two calls at
 * h266_inverse_trans.c:183-199: line j reads coefficient k at in[k*is + j]
 * and writes sample i at out[j*os + i]; the last skipLine lines are written
 * as zero without reading their input. */
void or_IDct1d(int type, int n, const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
               int32_t line, int32_t skipLine, int32_t shift)
{
    int32_t T[64 * 64];
    if (or_TrMatrix(type, n, T)) return;
    const int nz = kept_outputs(type, n);
    const int64_t rnd = shift > 0 ? ((int64_t)1 << (shift - 1)) : 0;
    for (int32_t j = 0; j < line; j++)
        for (int i = 0; i < n; i++) {
            if (j >= line - skipLine) { out[j * os + i] = 0; continue; }
            int64_t s = 0;
            for (int k = 0; k < nz; k++) s += (int64_t)T[k * n + i] * in[k * is + j];
            out[j * os + i] = clip16((s + rnd) >> shift);
        }
}

/* Xin266IDct2WxH (h266_inverse_trans.c:51-89): vertical then horizontal
 * DCT-II through tmp (height-strided), coefficient columns past 32 skipped. */
void or_IDct2WxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
                 uint32_t w, uint32_t h, or_coeff *tmp, int32_t s1, int32_t s2)
{
    const int32_t skipW = w > 32 ? (int32_t)w - 32 : 0;
    or_IDct1d(OR_DCT2, (int)h, in, is, tmp, h, (int32_t)w, skipW, s1);
    or_IDct1d(OR_DCT2, (int)w, tmp, h, out, os, (int32_t)h, 0, s2);
}

/* Xin266ISkipWxH (h266_inverse_trans.c:91-113): arithmetic right shift by
 * 15 - 8 - ((lgW + lgH) >> 1), truncated to int16 (the shift uses the 8-bit
 * constant whatever the internal bit depth). */
void or_ISkipWxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os, uint32_t w, uint32_t h)
{
    int lgW = 0, lgH = 0;
    while ((1u << lgW) < w) lgW++;
    while ((1u << lgH) < h) lgH++;
    const int shift = 15 - 8 - ((lgW + lgH) >> 1);
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) out[y * os + x] = (int16_t)(in[y * is + x] >> shift);
}

/* Xin266IDctWxH (h266_inverse_trans.c:115-229) for the two-sided cases:
 * shifts 7 and 20 - bitDepth (:145-146), the MTS zero-out skip (:147), the
 * implicit intra-luma DST-VII (:149-159), then 2-D DCT-II, skip or two 1-D
 * passes.  Returns -1 for the one-sided branches (:201-228, whose result the
 * reference leaves in its section scratch buffer) and unsupported kernels. */
int or_IDctWxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
               uint32_t w, uint32_t h, uint32_t mtsIdx, int isIntraLuma, int enableMts, int bitDepth)
{
    int32_t hor, ver;
    int lgW = 0, lgH = 0;
    while ((1u << lgW) < w) lgW++;
    while ((1u << lgH) < h) lgH++;
    if (mtsIdx > 5) return -1;
    or_MtsTypes(mtsIdx, &hor, &ver);
    const int32_t s1 = 6 + 1, s2 = 6 + 15 - 1 - bitDepth;
    const int32_t skipW = (hor != OR_DCT2 && w == 32) ? 16 : (w > 32 ? (int32_t)w - 32 : 0);
    if (isIntraLuma && enableMts && mtsIdx == 0) {
        if (w >= 4 && w <= 16) hor = OR_DST7;
        if (h >= 4 && h <= 16) ver = OR_DST7;
    }
    if (hor == OR_SKIP) {
        or_ISkipWxH(in, is, out, os, w, h);
        return 0;
    }
    if (lgW == 0 || lgH == 0) return -1;
    int32_t T[64 * 64];
    if (or_TrMatrix(hor, (int)w, T) || or_TrMatrix(ver, (int)h, T)) return -1;
    or_coeff *tmp = (or_coeff *)malloc(sizeof(or_coeff) * w * h);
    if (hor == OR_DCT2 && ver == OR_DCT2) {
        or_IDct2WxH(in, is, out, os, w, h, tmp, s1, s2);
    } else {
        or_IDct1d(ver, (int)h, in, is, tmp, h, (int32_t)w, skipW, s1);
        or_IDct1d(hor, (int)w, tmp, h, out, os, (int32_t)h, 0, s2);
    }
    free(tmp);
    return 0;
}

/* XinBlockAddForDct (h266_func_struct.h:530-538; Xin26xBlockRecon is
 * closed): recon = Clip3(0, 2^bitDepth - 1, pred + residual). */
void or_BlockRecon(const or_coeff *res, intptr_t rs, const or_pixel *pred, intptr_t ps,
                   or_pixel *recon, intptr_t os, uint32_t w, uint32_t h, int bitDepth)
{
    const int32_t maxv = (1 << bitDepth) - 1;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            int32_t v = (int32_t)pred[y * ps + x] + res[y * rs + x];
            recon[y * os + x] = (or_pixel)(v < 0 ? 0 : v > maxv ? maxv : v);
        }
}

/* Xin266IDctAdd (h266_idct_add.c:45-104) without the LMCS chroma residual
 * scaling: inverse transform into a 64-strided buffer, then the recon add. */
int or_IDctAdd(const or_coeff *in, intptr_t is, const or_pixel *pred, intptr_t ps, or_pixel *out,
               intptr_t os, uint32_t w, uint32_t h, uint32_t mtsIdx, int isIntraLuma, int enableMts,
               int bitDepth)
{
    or_coeff buf[64 * 64];
    if (w > 64 || h > 64) return -1;
    if (or_IDctWxH(in, is, buf, 64, w, h, mtsIdx, isIntraLuma, enableMts, bitDepth)) return -1;
    or_BlockRecon(buf, 64, pred, ps, out, os, w, h, bitDepth);
    return 0;
}

/* ---------------------------------------------------------------------------
 * DMVR: bilinear prediction (h266_core/source/h266_bilinear_filter.c, source),
 * the S16 SAD with row subsampling (closed Xin26xComputeSadS16,
 * h26x_core/include/h26x_compute_dist.h:942-950) and the decoder-side MV
 * refinement of Xin266DeriveMvFromDmvr (h266_encoder/source/
 * h266_inter_prediction.c:864-1195, source).  10-bit: BILINEAR_PREC equals the
 * internal bit depth (h266_inter_pred_context.h:42-44), so Copy keeps the
 * samples and both filter passes shift by 4 with offset 8.
 * ------------------------------------------------------------------------- */
#define OR_BILI_STRIDE 256      /* BILI_BUF_STRIDE = XIN_MAX_CU_SIZE * 2 */

/* Every entry writes width + 4 columns (the DMVR padding) of `height` rows. */
void or_BiliInterpCopy(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                       uint32_t w, uint32_t h)
{
    (void)frac;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w + 4; x++) dst[y * ds + x] = (int16_t)src[y * ss + x];
}

/* two taps (16 - f, f) on samples x and x + 1, (sum + 8) >> 4 */
static int16_t bili2(int32_t a, int32_t b, int32_t f) { return (int16_t)(((16 - f) * a + f * b + 8) >> 4); }

void or_BiliInterpHor(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                      uint32_t w, uint32_t h)
{
    const int32_t f = frac & 15;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w + 4; x++) dst[y * ds + x] = bili2(src[y * ss + x], src[y * ss + x + 1], f);
}

void or_BiliInterpVet(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                      uint32_t w, uint32_t h)
{
    const int32_t f = frac >> 4;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w + 4; x++) dst[y * ds + x] = bili2(src[y * ss + x], src[(y + 1) * ss + x], f);
}

/* horizontal pass over height + 1 rows into int16, then the vertical pass on it */
void or_BiliInterpHorVet(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                         uint32_t w, uint32_t h)
{
    const int32_t fx = frac & 15, fy = frac >> 4;
    int16_t *t = (int16_t *)malloc(sizeof(int16_t) * (w + 4) * (h + 1));
    for (uint32_t y = 0; y < h + 1; y++)
        for (uint32_t x = 0; x < w + 4; x++) t[y * (w + 4) + x] = bili2(src[y * ss + x], src[y * ss + x + 1], fx);
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w + 4; x++)
            dst[y * ds + x] = bili2(t[y * (w + 4) + x], t[(y + 1) * (w + 4) + x], fy);
    free(t);
}

/* SAD over the rows y = 0, 2^subShift, 2*2^subShift, ..., scaled back by
 * << subShift (the subsampled-SAD convention of HM/VTM's DistParam.subShift,
 * which the reference's DMVR cost follows with its >> 1). */
uint32_t or_ComputeSadS16(const int16_t *a, intptr_t as, const int16_t *b, intptr_t bs, uint32_t w, uint32_t h,
                          uint32_t subShift)
{
    uint32_t sad = 0;
    for (uint32_t y = 0; y < h; y += 1u << subShift)
        for (uint32_t x = 0; x < w; x++) {
            const int32_t d = (int32_t)a[y * as + x] - b[y * bs + x];
            sad += (uint32_t)(d < 0 ? -d : d);
        }
    return sad << subShift;
}

/* Xin266DivForMaxQ7 (:864-912): three-bit quotient of N / D in 1/16 units, clamped to 8. */
static int32_t div_q7(int64_t N, int64_t D)
{
    int sign = 0;
    int32_t q = 0;
    if (N < 0) { sign = 1; N = -N; }
    D <<= 3;
    if (N >= D) { N -= D; q++; }
    q <<= 1;
    D >>= 1;
    if (N >= D) { N -= D; q++; }
    q <<= 1;
    if (N >= (D >> 1)) q++;
    return sign ? -q : q;
}

/* Xin266SubPelRefineDmvr / Xin266SubPelErrorSrfc (:914-974): the parametric
 * error surface around the integer minimum c = costs[0] with neighbours
 * left, top, right, bottom = costs[1..4]. */
void or_DmvrSubPel(int non0Cost, int16_t *dmv, const int64_t costs[5])
{
    if (!non0Cost || abs(dmv[0]) == 32 || abs(dmv[1]) == 32) return;
    for (int hv = 0; hv < 2; hv++) {
        const int64_t num = (costs[hv + 1] - costs[hv + 3]) * 16;
        const int64_t den = costs[hv + 1] + costs[hv + 3] - 2 * costs[0];
        if (den == 0) continue;
        int32_t d;
        if (costs[hv + 1] != costs[0] && costs[hv + 3] != costs[0]) d = div_q7(num, den);
        else d = costs[hv + 1] == costs[0] ? -8 : 8;
        dmv[hv] = (int16_t)(dmv[hv] + d);
    }
}

/* The L0 or L1 bilinear prediction of a DMVR block: the clipped MV moved by
 * -2 pels in both directions, (w + 4) x (h + 4) int16 samples at stride 256
 * (:1033-1090). */
static void dmvr_bili(const or_pixel *ref, intptr_t rs, int32_t cuX, int32_t cuY, int32_t mvx, int32_t mvy,
                      const or_dmvr *d, int16_t *buf)
{
    mvx = mvx < d->minMvX * 16 ? d->minMvX * 16 : mvx > d->maxMvX * 16 ? d->maxMvX * 16 : mvx;
    mvy = mvy < d->minMvY * 16 ? d->minMvY * 16 : mvy > d->maxMvY * 16 ? d->maxMvY * 16 : mvy;
    const int32_t ox = cuX * 16 + mvx - 32, oy = cuY * 16 + mvy - 32;
    const int32_t fx = ox & 15, fy = oy & 15, frac = (fy << 4) + fx;
    const or_pixel *src = ref + (intptr_t)(oy >> 4) * rs + (ox >> 4);
    if (fx && fy) or_BiliInterpHorVet(src, rs, buf, OR_BILI_STRIDE, frac, d->width, d->height + 4);
    else if (fx) or_BiliInterpHor(src, rs, buf, OR_BILI_STRIDE, frac, d->width, d->height + 4);
    else if (fy) or_BiliInterpVet(src, rs, buf, OR_BILI_STRIDE, frac, d->width, d->height + 4);
    else or_BiliInterpCopy(src, rs, buf, OR_BILI_STRIDE, frac, d->width, d->height + 4);
}

/* Xin266DeriveMvFromDmvr (:976-1195): per subblock of min(16, side), an
 * integer search over the 25 offsets of the 5x5 window (mirrored between the
 * lists) with the early exit on a small centre cost, then the parametric
 * sub-pel step.  mvd receives one (x, y) delta in 1/16 pel per subblock,
 * raster order. */
void or_DeriveMvFromDmvr(const or_dmvr *d, int16_t *mvd)
{
    static const int8_t kSearch[25][2] = {
        {-2, -2}, {-1, -2}, {0, -2}, {1, -2}, {2, -2}, {-2, -1}, {-1, -1}, {0, -1}, {1, -1}, {2, -1},
        {-2, 0},  {-1, 0},  {0, 0},  {1, 0},  {2, 0},  {-2, 1},  {-1, 1},  {0, 1},  {1, 1},  {2, 1},
        {-2, 2},  {-1, 2},  {0, 2},  {1, 2},  {2, 2}};
    const uint32_t subW = d->width < 16 ? d->width : 16, subH = d->height < 16 ? d->height : 16;
    int16_t *p0 = (int16_t *)malloc(sizeof(int16_t) * OR_BILI_STRIDE * OR_BILI_STRIDE);
    int16_t *p1 = (int16_t *)malloc(sizeof(int16_t) * OR_BILI_STRIDE * OR_BILI_STRIDE);
    dmvr_bili(d->ref0, d->stride0, d->cuPelX, d->cuPelY, d->mv0x, d->mv0y, d, p0);
    dmvr_bili(d->ref1, d->stride1, d->cuPelX, d->cuPelY, d->mv1x, d->mv1y, d, p1);
    int n = 0;
    for (uint32_t y0 = 0; y0 < d->height; y0 += subH)
        for (uint32_t x0 = 0; x0 < d->width; x0 += subW) {
            uint32_t cost[25];
            for (int i = 0; i < 25; i++) cost[i] = UINT32_MAX;
            const int16_t *c0 = p0 + (2 + y0) * OR_BILI_STRIDE + x0 + 2;
            const int16_t *c1 = p1 + (2 + y0) * OR_BILI_STRIDE + x0 + 2;
            int non0 = 1;
            int best = 12;                          /* index of the minimum in the 5x5 grid */
            uint32_t minCost = or_ComputeSadS16(c0, OR_BILI_STRIDE, c1, OR_BILI_STRIDE, subW, subH, 1) >> 1;
            minCost -= minCost >> 2;
            if (minCost < subH * subW) {
                non0 = 0;
            } else {
                cost[12] = minCost;
                for (int i = 0; i < 25; i++) {
                    const int dx = kSearch[i][0], dy = kSearch[i][1];
                    if (cost[i] == UINT32_MAX)
                        cost[i] = or_ComputeSadS16(c0 + dy * OR_BILI_STRIDE + dx, OR_BILI_STRIDE,
                                                   c1 - dy * OR_BILI_STRIDE - dx, OR_BILI_STRIDE, subW, subH, 1) >> 1;
                    if (cost[i] < minCost) { minCost = cost[i]; best = i; }
                }
            }
            int16_t dmv[2] = {(int16_t)(kSearch[best][0] * 16), (int16_t)(kSearch[best][1] * 16)};
            if (non0) {
                const int64_t cs[5] = {cost[best], cost[best - 1 < 0 ? 0 : best - 1], cost[best - 5 < 0 ? 0 : best - 5],
                                       cost[best + 1 > 24 ? 24 : best + 1], cost[best + 5 > 24 ? 24 : best + 5]};
                or_DmvrSubPel(non0, dmv, cs);
            }
            mvd[2 * n] = dmv[0];
            mvd[2 * n + 1] = dmv[1];
            n++;
        }
    free(p0);
    free(p1);
}

/* Lookahead downscale (the role of the closed Xin26xDownscale2x2,
 * h266_encoder/include/h266_func_struct.h:425-433): each output sample is the
 * rounded mean of its 2x2 source block.  A design choice, not a restatement. */
void or_Downscale2x2(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds, uint32_t w, uint32_t h)
{
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const or_pixel *p = src + (intptr_t)(2 * y) * ss + 2 * x;
            dst[y * ds + x] = (or_pixel)((p[0] + p[1] + p[ss] + p[ss + 1] + 2) >> 2);
        }
}

/* ---------------------------------------------------------------------------
 * Frame-level drivers (checker for the batched kernels).
 * ------------------------------------------------------------------------- */

/* MV window of a CU, h266_encoder/source/h266_enc_init.c:1462-1465 with
 * XIN_PADDING_OFFSET_X/Y = 16 (h26x_definition.h:103-104). */
static void cu_mv_window(int32_t x, int32_t y, int32_t w, int32_t h, int32_t fw, int32_t fh,
                         int32_t ctu, int32_t *mnx, int32_t *mny, int32_t *mxx, int32_t *mxy)
{
    *mnx = (int16_t)(-x - ctu - (16 - 8));
    *mny = (int16_t)(-y - ctu - (16 - 8));
    *mxx = (int16_t)(fw + (ctu - w + 16 - 8) - x);
    *mxy = (int16_t)(fh + (ctu - h + 16 - 8) - y);
}

static inline int32_t clip3(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

void or_RefineFracBlocks(const or_pixel *curY, const or_pixel *curU, const or_pixel *curV,
                         intptr_t csY, intptr_t csC,
                         const or_pixel *refY, const or_pixel *refU, const or_pixel *refV,
                         intptr_t rsY, intptr_t rsC, int32_t fw, int32_t fh,
                         const or_block *blk, int32_t nblk, const or_refine_prm *prm,
                         uint32_t *candCost, int32_t *bestIdx, uint32_t *bestCost,
                         int32_t *bestMvx, int32_t *bestMvy,
                         or_pixel *predY, or_pixel *predU, or_pixel *predV,
                         intptr_t psY, intptr_t psC)
{
    or_pixel *bufY = (or_pixel *)malloc(sizeof(or_pixel) * 64 * 64);
    or_pixel *bufU = (or_pixel *)malloc(sizeof(or_pixel) * 32 * 32);
    or_pixel *bufV = (or_pixel *)malloc(sizeof(or_pixel) * 32 * 32);
    for (int32_t b = 0; b < nblk; b++) {
        const or_block *B = &blk[b];
        int32_t mnx, mny, mxx, mxy;
        cu_mv_window(B->x, B->y, B->w, B->h, fw, fh, prm->ctuSize, &mnx, &mny, &mxx, &mxy);
        uint32_t best = UINT32_MAX;
        int32_t bi = 0, bmx = 0, bmy = 0;
        for (int32_t pass = 0; pass < 2; pass++) {
            /* pass 0 costs every candidate; pass 1 writes the MC of the best. */
            int32_t c0 = pass == 0 ? 0 : bi, c1 = pass == 0 ? prm->ncand : bi + 1;
            for (int32_t c = c0; c < c1; c++) {
                /* MV clip and address split: h266_inter_prediction.c:332-342 (luma),
                 * :664-674 (chroma, 1/32 of chroma sample). */
                int32_t mvx = clip3(B->mvx + prm->offx[c], mnx * 16, mxx * 16);
                int32_t mvy = clip3(B->mvy + prm->offy[c], mny * 16, mxy * 16);
                int32_t rx = (B->x << 4) + mvx, ry = (B->y << 4) + mvy;
                int32_t fracY = ((ry & 15) << 4) + (rx & 15);
                or_pixel *dY = pass ? predY + B->y * psY + B->x : bufY;
                intptr_t sdY = pass ? psY : 64;
                or_LumaInterpolation(refY + (ry >> 4) * rsY + (rx >> 4), rsY, dY, sdY, fracY,
                                     (uint32_t)B->w, (uint32_t)B->h, 0);
                uint32_t satd = 0;
                if (pass == 0)
                    satd = or_ComputeSatd(curY + B->y * csY + B->x, csY, bufY, 64, B->w, B->h);
                if (prm->useChroma || pass) {
                    int32_t fracC = ((ry & 31) << 5) + (rx & 31);
                    intptr_t off = (ry >> 5) * rsC + (rx >> 5);
                    int32_t cw = B->w / 2, ch = B->h / 2;
                    or_pixel *dU = pass ? predU + (B->y / 2) * psC + B->x / 2 : bufU;
                    or_pixel *dV = pass ? predV + (B->y / 2) * psC + B->x / 2 : bufV;
                    intptr_t sdC = pass ? psC : 32;
                    or_ChromaInterpolation(refU + off, rsC, dU, sdC, fracC, cw, ch);
                    or_ChromaInterpolation(refV + off, rsC, dV, sdC, fracC, cw, ch);
                    if (pass == 0) {
                        intptr_t co = (B->y / 2) * csC + B->x / 2;
                        satd += or_ComputeSatd(curU + co, csC, bufU, 32, cw, ch);
                        satd += or_ComputeSatd(curV + co, csC, bufV, 32, cw, ch);
                    }
                }
                if (pass == 0) {
                    uint32_t cost = (satd << 8) + or_MvdBits(mvx - B->mvpx, mvy - B->mvpy) * prm->lambda;
                    candCost[(size_t)b * prm->ncand + c] = cost;
                    if (cost < best) { best = cost; bi = c; bmx = mvx; bmy = mvy; }
                }
            }
        }
        bestIdx[b] = bi;
        bestCost[b] = best;
        bestMvx[b] = bmx;
        bestMvy[b] = bmy;
    }
    free(bufY);
    free(bufU);
    free(bufV);
}

/* The integer search of the serving path for blocks [i0, i1) of blk: per block
 * the window +/-range around the predictor (1/16 pel, taken to quarter pel by
 * >> 2 and rounded to integer pel as Xin266FillMeCand rounds,
 * h266_analyze_inter_cu.c:271-272) with that centre first clipped into the CU
 * MV window (h266_enc_init.c:1462-1465), then the window intersected with it;
 * raster scan, cost (SAD << 8) + bits(4 mv - mvp / 4) * sadLambda, strict <
 * keeps the earliest (the rule of or_MotionEstFull).  The winner goes to the
 * block's mvx/mvy in 1/16 pel.  cur/ref point at sample (0,0) of their planes
 * (ref padded). */
void or_SearchBlocks(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs, int32_t fw, int32_t fh,
                     or_block *blk, int32_t i0, int32_t i1, int32_t range, uint32_t sadLambda, int32_t ctuSize)
{
    for (int32_t i = i0; i < i1; i++) {
        or_block *B = &blk[i];
        int32_t mnx, mny, mxx, mxy;
        cu_mv_window(B->x, B->y, B->w, B->h, fw, fh, ctuSize, &mnx, &mny, &mxx, &mxy);
        const int32_t pqx = B->mvpx >> 2, pqy = B->mvpy >> 2;
        const int32_t cx = clip3((pqx + 2) >> 2, mnx, mxx), cy = clip3((pqy + 2) >> 2, mny, mxy);
        const int32_t x0 = cx - range > mnx ? cx - range : mnx, x1 = cx + range < mxx ? cx + range : mxx;
        const int32_t y0 = cy - range > mny ? cy - range : mny, y1 = cy + range < mxy ? cy + range : mxy;
        uint32_t best = UINT32_MAX;
        int32_t bx = cx, by = cy;
        for (int32_t my = y0; my <= y1; my++)
            for (int32_t mx = x0; mx <= x1; mx++) {
                const uint32_t sad = or_ComputeSad(cur + B->y * cs + B->x, cs, ref + (B->y + my) * rs + B->x + mx,
                                                   rs, (uint32_t)B->w, (uint32_t)B->h);
                const uint32_t c = (sad << 8) + or_MvdBits(4 * mx - pqx, 4 * my - pqy) * sadLambda;
                if (c < best) { best = c; bx = mx; by = my; }
            }
        B->mvx = 16 * bx;
        B->mvy = 16 * by;
    }
}

/* ---- AMVP predictor selection ---------------------------------------- */
/* h266_get_neighbour_mv.c:212-243: left shift when the precision rises,
 * rounding right shift (half toward zero) when it drops. */
int32_t or_ChangeMvPrec(int32_t v, int32_t inPrec, int32_t outPrec)
{
    return or_MvRoundShift(v, inPrec - outPrec);
}

/* amvrPrecision[] of h266_entropy_write.c:120-123, indexed by imvIdx. */
static const int32_t kAmvrPrec[4] = {4, 2, 0, 3};

static int field_mv(const int32_t *field, intptr_t stride, int32_t fw, int32_t fh, int32_t px, int32_t py,
                    int32_t out[2])
{
    if (px < 0 || py < 0 || px >= fw || py >= fh) return 0;
    const int32_t *u = field + ((intptr_t)(py >> 2) * stride + (px >> 2)) * 2;
    out[0] = u[0];
    out[1] = u[1];
    return 1;
}

/* h266_get_neighbour_mv.c:2451-2560: the first available of the left group
 * (below-left A0, then left A1), the first of the top group (above-right B0,
 * above B1, above-left B2), a duplicate pair reduced to one, a zero MV added
 * when the list is empty or its one entry is non-zero, two candidates kept
 * only when they differ. */
int32_t or_AmvpFromField(const int32_t *field, intptr_t stride, int32_t fw, int32_t fh, int32_t x, int32_t y,
                         int32_t w, int32_t h, int32_t cand[4])
{
    int32_t list[3][2];
    int n = 0;
    const int32_t lx[2] = {x - 1, x - 1}, ly[2] = {y + h, y + h - 1};
    for (int k = 0; k < 2; k++)
        if (field_mv(field, stride, fw, fh, lx[k], ly[k], list[n])) { n++; break; }
    const int32_t tx[3] = {x + w, x + w - 1, x - 1}, ty[3] = {y - 1, y - 1, y - 1};
    for (int k = 0; k < 3; k++)
        if (field_mv(field, stride, fw, fh, tx[k], ty[k], list[n])) { n++; break; }
    if (n == 2 && list[0][0] == list[1][0] && list[0][1] == list[1][1]) n = 1;
    if (n < 1 || (n == 1 && (list[0][0] != 0 || list[0][1] != 0))) {
        list[n][0] = list[n][1] = 0;
        n++;
    }
    const int two = n >= 2 && (list[0][0] != list[1][0] || list[0][1] != list[1][1]);
    cand[0] = list[0][0];
    cand[1] = list[0][1];
    cand[2] = two ? list[1][0] : 0;
    cand[3] = two ? list[1][1] : 0;
    return two ? 2 : 1;
}

/* h266_analyze_inter_cu.c:60-131 with the rounding at the end of
 * Xin266FillAmvpCand (h266_get_neighbour_mv.c:2549-2558) and the luma
 * prediction of Xin266LumaInterPred (h266_inter_prediction.c:302-354). */
int32_t or_SelectMvp(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs, int32_t fw, int32_t fh,
                     int32_t x, int32_t y, int32_t w, int32_t h, int32_t ncand, const int32_t *cand,
                     int32_t imvIdx, int32_t ctuSize, uint32_t *sad, int32_t mv[2], int16_t meMv[2])
{
    int32_t mnx, mny, mxx, mxy;
    cu_mv_window(x, y, w, h, fw, fh, ctuSize, &mnx, &mny, &mxx, &mxy);
    const int32_t prec = kAmvrPrec[imvIdx & 3];
    or_pixel *pred = (or_pixel *)malloc(sizeof(or_pixel) * (size_t)w * (size_t)h);
    uint32_t best = UINT32_MAX;
    int32_t idx = 0, rounded[2][2] = {{0, 0}, {0, 0}};
    for (int32_t c = 0; c < ncand && c < 2; c++) {
        for (int k = 0; k < 2; k++)
            rounded[c][k] = or_ChangeMvPrec(or_ChangeMvPrec(cand[2 * c + k], 6, prec), prec, 6);
        const int32_t mx = clip3(rounded[c][0], mnx * 16, mxx * 16), my = clip3(rounded[c][1], mny * 16, mxy * 16);
        const int32_t ox = x * 16 + mx, oy = y * 16 + my;
        const int32_t frac = ((oy & 15) << 4) + (ox & 15);
        or_LumaInterpolation(ref + (intptr_t)(oy >> 4) * rs + (ox >> 4), rs, pred, w, frac, (uint32_t)w, (uint32_t)h,
                             imvIdx == 3 ? 1u : 0u);
        const uint32_t s = or_ComputeSad(cur + (intptr_t)y * cs + x, cs, pred, w, (uint32_t)w, (uint32_t)h);
        if (s < best) {
            best = s;
            idx = c;
        }
    }
    free(pred);
    *sad = best;
    mv[0] = rounded[idx][0];
    mv[1] = rounded[idx][1];
    meMv[0] = (int16_t)or_ChangeMvPrec(rounded[idx][0], 6, 4);
    meMv[1] = (int16_t)or_ChangeMvPrec(rounded[idx][1], 6, 4);
    return idx;
}

/* ---- MCTF motion estimation ----------------------------------------------
 * The reference's MCTF motion search (Xin26xMctfMotionEstFrame,
 * h26x_core/include/h26x_mctf.h:27, per level and block size; its subsampling
 * Xin26xMctfSubSample, :35) is closed.  This is the library's own design in
 * that role, hierarchical as the published temporal filter of VVC's
 * reference software is: 2x2-mean subsampling, candidates from the lower
 * level, an integer search, fractional rounds at the finest level; the
 * fractional prediction is the standard's 8-tap luma filter. */
void or_MctfSubsample(const or_pixel *src, intptr_t ss, int32_t w, int32_t h, or_pixel *dst, intptr_t ds,
                      int32_t pad)
{
    const int32_t hw = w / 2, hh = h / 2;
    for (int32_t y = -pad; y < hh + pad; y++)
        for (int32_t x = -pad; x < hw + pad; x++) {
            const int32_t cx = clip3(x, 0, hw - 1), cy = clip3(y, 0, hh - 1);
            const or_pixel *p = src + (intptr_t)(2 * cy) * ss + 2 * cx;
            dst[(intptr_t)y * ds + x] = (or_pixel)((p[0] + p[1] + p[ss] + p[ss + 1] + 2) >> 2);
        }
}

/* SSD of the bs x bs block at (x, y) against the reference at MV (mx, my)
 * (1/16 pel); UINT32_MAX when the 8-tap reach leaves the padding. */
static uint32_t mctf_ssd(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs, int32_t w, int32_t h,
                         int32_t pad, int32_t bs, int32_t x, int32_t y, int32_t mx, int32_t my, or_pixel *buf)
{
    const int32_t px = x + (mx >> 4), py = y + (my >> 4);
    if (px - 3 < -pad || py - 3 < -pad || px + bs + 4 > w - 1 + pad || py + bs + 4 > h - 1 + pad) return UINT32_MAX;
    or_LumaInterpolation(ref + (intptr_t)py * rs + px, rs, buf, bs, ((my & 15) << 4) + (mx & 15), (uint32_t)bs,
                         (uint32_t)bs, 0);
    uint32_t e = 0;
    for (int32_t r = 0; r < bs; r++)
        for (int32_t c = 0; c < bs; c++) {
            const int32_t d = (int32_t)cur[(intptr_t)(y + r) * cs + x + c] - (int32_t)buf[r * bs + c];
            e += (uint32_t)(d * d);
        }
    return e;
}

void or_MctfMeLevel(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs, int32_t w, int32_t h,
                    int32_t pad, int32_t bs, int32_t range, const int32_t *lower, int32_t lowerBw, int32_t lowerBh,
                    int32_t factor, int32_t frac, int32_t by0, int32_t by1, int32_t *mv, uint32_t *err)
{
    const int32_t nbx = w / bs, nby = h / bs;
    or_pixel *buf = (or_pixel *)malloc(sizeof(or_pixel) * (size_t)bs * (size_t)bs);
    for (int32_t by = by0 < 0 ? 0 : by0; by < by1 && by < nby; by++)
        for (int32_t bx = 0; bx < nbx; bx++) {
            const int32_t x = bx * bs, y = by * bs;
            uint32_t best = UINT32_MAX;
            int32_t bmx = 0, bmy = 0;
            if (lower) {
                for (int32_t py = -1; py <= 1; py++)
                    for (int32_t px = -1; px <= 1; px++) {
                        const int32_t tx = bx / 2 + px, ty = by / 2 + py;
                        if (tx < 0 || ty < 0 || tx >= lowerBw || ty >= lowerBh) continue;
                        const int32_t cx = lower[2 * ((intptr_t)ty * lowerBw + tx)] * factor;
                        const int32_t cy = lower[2 * ((intptr_t)ty * lowerBw + tx) + 1] * factor;
                        const uint32_t e = mctf_ssd(cur, cs, ref, rs, w, h, pad, bs, x, y, cx, cy, buf);
                        if (e < best) { best = e; bmx = cx; bmy = cy; }
                    }
                const uint32_t e = mctf_ssd(cur, cs, ref, rs, w, h, pad, bs, x, y, 0, 0, buf);
                if (e < best) { best = e; bmx = 0; bmy = 0; }
            }
            const int32_t ix = bmx / 16, iy = bmy / 16;
            for (int32_t y2 = iy - range; y2 <= iy + range; y2++)
                for (int32_t x2 = ix - range; x2 <= ix + range; x2++) {
                    const uint32_t e = mctf_ssd(cur, cs, ref, rs, w, h, pad, bs, x, y, 16 * x2, 16 * y2, buf);
                    if (e < best) { best = e; bmx = 16 * x2; bmy = 16 * y2; }
                }
            if (frac) {
                for (int32_t step = 4; step >= 1; step -= 3) {
                    const int32_t ox = bmx, oy = bmy;
                    for (int32_t y2 = -3 * step; y2 <= 3 * step; y2 += step)
                        for (int32_t x2 = -3 * step; x2 <= 3 * step; x2 += step) {
                            const uint32_t e = mctf_ssd(cur, cs, ref, rs, w, h, pad, bs, x, y, ox + x2, oy + y2, buf);
                            if (e < best) { best = e; bmx = ox + x2; bmy = oy + y2; }
                        }
                }
            }
            const intptr_t b = (intptr_t)by * nbx + bx;
            mv[2 * b] = bmx;
            mv[2 * b + 1] = bmy;
            err[b] = best;
        }
    free(buf);
}

void or_FullSearchFrame(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs,
                        int32_t fw, int32_t fh, int32_t bw, int32_t bh, int32_t range,
                        uint32_t sadLambda, int32_t ctuSize, int32_t blk0, int32_t blk1,
                        int16_t *bestMvX, int16_t *bestMvY, uint32_t *bestCost, uint32_t *bestSad)
{
    const int32_t nbx = fw / bw;
    for (int32_t b = blk0; b < blk1; b++) {
        int32_t x = (b % nbx) * bw, y = (b / nbx) * bh;
        int32_t mnx, mny, mxx, mxy;
        cu_mv_window(x, y, bw, bh, fw, fh, ctuSize, &mnx, &mny, &mxx, &mxy);
        or_me me;
        memset(&me, 0, sizeof(me));
        me.width = (uint32_t)bw;
        me.height = (uint32_t)bh;
        me.minMvX = (int16_t)mnx; me.minMvY = (int16_t)mny;
        me.maxMvX = (int16_t)mxx; me.maxMvY = (int16_t)mxy;
        me.searchWidth = range;
        me.searchHeight = range;
        me.sadLambda = sadLambda;
        me.input = cur + y * cs + x;
        me.inputStride = cs;
        me.ref = ref + y * rs + x;
        me.refStride = rs;
        or_MotionEstFull(&me);
        bestMvX[b] = me.bestMvX;
        bestMvY[b] = me.bestMvY;
        bestCost[b] = me.bestCost;
        bestSad[b] = me.bestSad;
    }
}

/* Forward transform of a batch of residual blocks laid out as the batched
 * GPU entry takes them (xg_fdct_plan): block i is int16 at resid + offset[i],
 * row-major with stride = width, desc[i] = lgW | lgH << 4 | mtsIdx << 8;
 * coefficients go to coeff + offset[i], stride = width.  Each block is
 * or_FDctWxH (h266_forward_trans.c:45-152) with the explicit mtsIdx, 10-bit. */
void or_FDctBatch(const int16_t *resid, const int64_t *offset, const uint32_t *desc, int32_t i0, int32_t i1,
                  or_coeff *coeff)
{
    or_coeff tmp[64 * 64];
    for (int32_t i = i0; i < i1; i++) {
        const uint32_t w = 1u << (desc[i] & 15), h = 1u << ((desc[i] >> 4) & 15), mts = (desc[i] >> 8) & 15;
        const int16_t *r = resid + offset[i];
        for (uint32_t k = 0; k < w * h; k++) tmp[k] = r[k];
        or_FDctWxH(tmp, w, coeff + offset[i], w, w, h, mts, 0, 0, 10);
    }
}

/* Blocks [i0, i1) of a reconstruction batch in the layout of the batched GPU
 * entry: int16 coefficients, prediction and recon at offset[i], desc = lgW |
 * lgH << 4 | mtsIdx << 8 | implicitMts << 12; per block or_IDctAdd
 * (h266_idct_add.c:45-104) with 10-bit samples. */
void or_ReconBatch(const int16_t *coeff, const or_pixel *pred, const int64_t *offset, const uint32_t *desc,
                   int32_t i0, int32_t i1, or_pixel *recon)
{
    or_coeff tmp[64 * 64];
    for (int32_t i = i0; i < i1; i++) {
        const uint32_t w = 1u << (desc[i] & 15), h = 1u << ((desc[i] >> 4) & 15), mts = (desc[i] >> 8) & 15;
        const int implicit = (desc[i] >> 12) & 1;
        const int16_t *c = coeff + offset[i];
        for (uint32_t k = 0; k < w * h; k++) tmp[k] = c[k];
        or_IDctAdd(tmp, w, pred + offset[i], w, recon + offset[i], w, w, h, mts, implicit, implicit, 10);
    }
}

static int32_t median3(int32_t a, int32_t b, int32_t c)
{
    int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return c < lo ? lo : (c > hi ? hi : c);
}

/* Stage 2 inputs of the pipeline for block b and reference r, from the
 * stage-1 full-search MVs fullMv[(b*2+r)*2 + {0,1}] (integer pel) of a raster
 * grid nbx blocks wide: the predictor is the component-wise median of the
 * left, top and top-right blocks' MVs in quarter pel (missing neighbours
 * count as (0,0)); the start candidates are Xin266FillMeCand of that predictor
 * and the own, left, top and top-right MVs in that order
 * (h266_analyze_inter_cu.c:133-293: (0,0), predictor, then the parent and
 * neighbour MVs, each rounded to integer pel and de-duplicated). */
void or_PipelineCands(const int16_t *fullMv, int32_t nbx, int32_t b, int32_t r, int16_t *predX, int16_t *predY,
                      int16_t *candX, int16_t *candY, uint32_t *candNum)
{
    const int32_t bx = b % nbx, by = b / nbx;
    int16_t nq[4][2];                       /* own, left, top, top-right, quarter pel */
    const int32_t nbi[4] = {b, bx > 0 ? b - 1 : -1, by > 0 ? b - nbx : -1,
                            (by > 0 && bx + 1 < nbx) ? b - nbx + 1 : -1};
    for (int k = 0; k < 4; k++)
        for (int c = 0; c < 2; c++)
            nq[k][c] = nbi[k] < 0 ? 0 : (int16_t)(fullMv[(nbi[k] * 2 + r) * 2 + c] * 4);
    const int16_t px = (int16_t)median3(nq[1][0], nq[2][0], nq[3][0]);
    const int16_t py = (int16_t)median3(nq[1][1], nq[2][1], nq[3][1]);
    const int16_t ex[4] = {nq[0][0], nq[1][0], nq[2][0], nq[3][0]};
    const int16_t ey[4] = {nq[0][1], nq[1][1], nq[2][1], nq[3][1]};
    or_FillMeCand(px, py, ex, ey, 4, candX, candY, candNum);
    *predX = px;
    *predY = py;
}

/* Full motion-search pipeline over a raster grid of bw x bh blocks with two
 * reference pictures (a build composition of the reference's stages):
 *  1. full search +/-range around (0,0), predictor (0,0), per block and
 *     reference (or_MotionEstFull = closed Xin26xMotionEstFull);
 *  2. predictor = component-wise median of the pass-1 MVs of the left, top and
 *     top-right blocks (missing neighbours count as (0,0)), in quarter pel;
 *     start candidates = Xin266FillMeCand(predictor; own, left, top, top-right
 *     pass-1 MVs) (h266_analyze_inter_cu.c:133-293); diamond search +/-range
 *     around the rounded predictor (or_MotionEstDia = closed Xin26xMotionEstDia);
 *  3. half- then quarter-pel ring from the diamond result
 *     (h266_sub_pel_refine.c:81-212, or_FullSubPelMe);
 *  4. reference with the lower quarter-pel cost (ties: reference 0);
 *  5. 1/16-pel refinement + MC of the chosen reference (or_RefineFracBlocks),
 *     base MV = quarter-pel MV, predictor = that reference's predictor.
 * Pictures point at sample (0,0) of padded planes (luma stride sy, chroma sc).
 * Per block b and reference r: fullMv/intMv/qMv[(b*2+r)*2 + {0,1}] (int pel,
 * int pel, quarter pel), predMv likewise (quarter pel), intCost/qCost[b*2+r];
 * refIdx[b]; candCost[b*ncand + c], bestIdx/bestCost/bestMvx/bestMvy[b]. */
/* The pictures and parameters every stage of the pipeline reads. */
typedef struct {
    const or_pixel *cur[3];
    const or_pixel *ref[2][3];
    intptr_t csy, csc, rsy, rsc;
    int32_t fw, fh, bw, bh, nbx, range, ctu;
    uint32_t lambda;
    const or_refine_prm *prm;
} pipe_ctx;

/* Stage 1 of block b against reference r: full search +/-range around (0,0),
 * predictor (0,0); out = integer-pel MV. */
static void pipe_full(const pipe_ctx *c, int32_t b, int32_t r, int16_t out[2])
{
    const int32_t x = (b % c->nbx) * c->bw, y = (b / c->nbx) * c->bh;
    int32_t mnx, mny, mxx, mxy;
    cu_mv_window(x, y, c->bw, c->bh, c->fw, c->fh, c->ctu, &mnx, &mny, &mxx, &mxy);
    or_me me;
    memset(&me, 0, sizeof(me));
    me.width = (uint32_t)c->bw; me.height = (uint32_t)c->bh;
    me.minMvX = (int16_t)mnx; me.minMvY = (int16_t)mny; me.maxMvX = (int16_t)mxx; me.maxMvY = (int16_t)mxy;
    me.searchWidth = me.searchHeight = c->range;
    me.sadLambda = c->lambda;
    me.input = c->cur[0] + y * c->csy + x; me.inputStride = c->csy;
    me.ref = c->ref[r][0] + y * c->rsy + x; me.refStride = c->rsy;
    or_MotionEstFull(&me);
    out[0] = (int16_t)(me.bestMvX >> 2);
    out[1] = (int16_t)(me.bestMvY >> 2);
}

/* Stages 2-3 of block b against reference r, given the stage-1 MVs of b and
 * its left, top and top-right neighbours in the raster array fullMv:
 * predictor and start list (or_PipelineCands), diamond search, half- and
 * quarter-pel rings.  Outputs for this (b, r) at pred[2], imv[2], *icost,
 * qmv[2], *qcost. */
static void pipe_stage2(const pipe_ctx *c, const int16_t *fullMv, int32_t b, int32_t r, int16_t pred[2],
                        int16_t imv[2], uint32_t *icost, int16_t qmv[2], uint32_t *qcost)
{
    const int32_t x = (b % c->nbx) * c->bw, y = (b / c->nbx) * c->bh;
    int32_t mnx, mny, mxx, mxy;
    cu_mv_window(x, y, c->bw, c->bh, c->fw, c->fh, c->ctu, &mnx, &mny, &mxx, &mxy);
    or_me me;
    memset(&me, 0, sizeof(me));
    int16_t px, py;
    or_PipelineCands(fullMv, c->nbx, b, r, &px, &py, me.candX, me.candY, &me.candNum);
    pred[0] = px;
    pred[1] = py;
    me.width = (uint32_t)c->bw; me.height = (uint32_t)c->bh;
    me.minMvX = (int16_t)mnx; me.minMvY = (int16_t)mny; me.maxMvX = (int16_t)mxx; me.maxMvY = (int16_t)mxy;
    me.predMvX = px; me.predMvY = py;
    me.searchWidth = me.searchHeight = c->range;
    me.sadLambda = c->lambda;
    me.input = c->cur[0] + y * c->csy + x; me.inputStride = c->csy;
    me.ref = c->ref[r][0] + y * c->rsy + x; me.refStride = c->rsy;
    or_MotionEstDia(&me);
    imv[0] = (int16_t)(me.bestMvX >> 2);
    imv[1] = (int16_t)(me.bestMvY >> 2);
    *icost = me.bestCost;
    or_FullSubPelMe(&me);
    qmv[0] = me.bestMvX;
    qmv[1] = me.bestMvY;
    *qcost = me.bestCost;
}

/* Stages 4-5 of block b: the reference with the lower quarter-pel cost (ties:
 * reference 0), then its 1/16-pel refinement + MC around 4x the quarter-pel MV
 * with predictor 4x that reference's median predictor.  pred2/qmv2/qcost2 hold
 * both references' stage-2/3 results ([r*2 + c] and [r]). */
static int32_t pipe_final(const pipe_ctx *c, int32_t b, const int16_t *pred2, const int16_t *qmv2,
                          const uint32_t *qcost2, uint32_t *candCost, int32_t *bestIdx, uint32_t *bestCost,
                          int32_t *bestMvx, int32_t *bestMvy, or_pixel *predY, or_pixel *predU, or_pixel *predV,
                          intptr_t psy, intptr_t psc)
{
    const int32_t r = qcost2[1] < qcost2[0] ? 1 : 0;
    or_block blk;
    blk.x = (b % c->nbx) * c->bw; blk.y = (b / c->nbx) * c->bh; blk.w = c->bw; blk.h = c->bh;
    blk.mvx = qmv2[r * 2] * 4; blk.mvy = qmv2[r * 2 + 1] * 4;
    blk.mvpx = pred2[r * 2] * 4; blk.mvpy = pred2[r * 2 + 1] * 4;
    or_RefineFracBlocks(c->cur[0], c->cur[1], c->cur[2], c->csy, c->csc, c->ref[r][0], c->ref[r][1], c->ref[r][2],
                        c->rsy, c->rsc, c->fw, c->fh, &blk, 1, c->prm, candCost, bestIdx, bestCost, bestMvx, bestMvy,
                        predY, predU, predV, psy, psc);
    return r;
}

static pipe_ctx pipe_ctx_make(const or_pixel *curY, const or_pixel *curU, const or_pixel *curV, intptr_t csy,
                              intptr_t csc, const or_pixel *r0Y, const or_pixel *r0U, const or_pixel *r0V,
                              const or_pixel *r1Y, const or_pixel *r1U, const or_pixel *r1V, intptr_t rsy,
                              intptr_t rsc, int32_t fw, int32_t fh, int32_t bw, int32_t bh, int32_t range,
                              uint32_t sadLambda, int32_t ctuSize, const or_refine_prm *prm)
{
    pipe_ctx c;
    c.cur[0] = curY; c.cur[1] = curU; c.cur[2] = curV;
    c.ref[0][0] = r0Y; c.ref[0][1] = r0U; c.ref[0][2] = r0V;
    c.ref[1][0] = r1Y; c.ref[1][1] = r1U; c.ref[1][2] = r1V;
    c.csy = csy; c.csc = csc; c.rsy = rsy; c.rsc = rsc;
    c.fw = fw; c.fh = fh; c.bw = bw; c.bh = bh; c.nbx = fw / bw; c.range = range; c.ctu = ctuSize;
    c.lambda = sadLambda;
    c.prm = prm;
    return c;
}

void or_MePipeline(const or_pixel *curY, const or_pixel *curU, const or_pixel *curV, intptr_t csy, intptr_t csc,
                   const or_pixel *r0Y, const or_pixel *r0U, const or_pixel *r0V,
                   const or_pixel *r1Y, const or_pixel *r1U, const or_pixel *r1V, intptr_t rsy, intptr_t rsc,
                   int32_t fw, int32_t fh, int32_t bw, int32_t bh, int32_t range, uint32_t sadLambda,
                   int32_t ctuSize, const or_refine_prm *prm,
                   int16_t *fullMv, int16_t *predMv, int16_t *intMv, uint32_t *intCost, int16_t *qMv,
                   uint32_t *qCost, int32_t *refIdx, uint32_t *candCost, int32_t *bestIdx, uint32_t *bestCost,
                   int32_t *bestMvx, int32_t *bestMvy, or_pixel *predY, or_pixel *predU, or_pixel *predV,
                   intptr_t psy, intptr_t psc)
{
    const pipe_ctx c = pipe_ctx_make(curY, curU, curV, csy, csc, r0Y, r0U, r0V, r1Y, r1U, r1V, rsy, rsc, fw, fh,
                                     bw, bh, range, sadLambda, ctuSize, prm);
    const int32_t nb = c.nbx * (fh / bh);
    for (int32_t r = 0; r < 2; r++)
        for (int32_t b = 0; b < nb; b++) pipe_full(&c, b, r, fullMv + (b * 2 + r) * 2);
    for (int32_t r = 0; r < 2; r++)
        for (int32_t b = 0; b < nb; b++)
            pipe_stage2(&c, fullMv, b, r, predMv + (b * 2 + r) * 2, intMv + (b * 2 + r) * 2, intCost + b * 2 + r,
                        qMv + (b * 2 + r) * 2, qCost + b * 2 + r);
    for (int32_t b = 0; b < nb; b++)
        refIdx[b] = pipe_final(&c, b, predMv + b * 4, qMv + b * 4, qCost + b * 2,
                               candCost + (size_t)b * prm->ncand, bestIdx + b, bestCost + b, bestMvx + b,
                               bestMvy + b, predY, predU, predV, psy, psc);
}

/* The pipeline for the blocks sample[0..n) only: each needs the stage-1 MVs
 * of itself and its left, top and top-right neighbours, and nothing else, so
 * every output equals or_MePipeline's for that block.  Outputs are indexed by
 * sample position i (fullMv/predMv/intMv/qMv[(i*2+r)*2 + c], intCost/qCost
 * [i*2+r], refIdx/bestIdx/bestCost/bestMvx/bestMvy[i], candCost[i*ncand+k]);
 * the prediction of each sampled block is written into the planes at its
 * position.  Stage-1 results are computed once per block within a call, so a
 * sample of whole block rows costs one extra row of full searches.  Samples
 * may be split over threads (no shared state). */
void or_MePipelineSample(const or_pixel *curY, const or_pixel *curU, const or_pixel *curV, intptr_t csy,
                         intptr_t csc, const or_pixel *r0Y, const or_pixel *r0U, const or_pixel *r0V,
                         const or_pixel *r1Y, const or_pixel *r1U, const or_pixel *r1V, intptr_t rsy, intptr_t rsc,
                         int32_t fw, int32_t fh, int32_t bw, int32_t bh, int32_t range, uint32_t sadLambda,
                         int32_t ctuSize, const or_refine_prm *prm, const int32_t *sample, int32_t n,
                         int16_t *fullMv, int16_t *predMv, int16_t *intMv, uint32_t *intCost, int16_t *qMv,
                         uint32_t *qCost, int32_t *refIdx, uint32_t *candCost, int32_t *bestIdx, uint32_t *bestCost,
                         int32_t *bestMvx, int32_t *bestMvy, or_pixel *predY, or_pixel *predU, or_pixel *predV,
                         intptr_t psy, intptr_t psc)
{
    const pipe_ctx c = pipe_ctx_make(curY, curU, curV, csy, csc, r0Y, r0U, r0V, r1Y, r1U, r1V, rsy, rsc, fw, fh,
                                     bw, bh, range, sadLambda, ctuSize, prm);
    const int32_t nbx = c.nbx, nb = nbx * (fh / bh);
    int16_t *raster = (int16_t *)calloc((size_t)nb * 4, sizeof(int16_t));
    uint8_t *have = (uint8_t *)calloc((size_t)nb, 1);     /* stage 1 done for this block, this call */
    for (int32_t i = 0; i < n; i++) {
        const int32_t b = sample[i], bx = b % nbx, by = b / nbx;
        const int32_t nbi[4] = {b, bx > 0 ? b - 1 : -1, by > 0 ? b - nbx : -1,
                                (by > 0 && bx + 1 < nbx) ? b - nbx + 1 : -1};
        for (int k = 0; k < 4; k++)
            if (nbi[k] >= 0 && !have[nbi[k]]) {
                for (int32_t r = 0; r < 2; r++) pipe_full(&c, nbi[k], r, raster + (nbi[k] * 2 + r) * 2);
                have[nbi[k]] = 1;
            }
        for (int32_t r = 0; r < 2; r++) {
            fullMv[(i * 2 + r) * 2] = raster[(b * 2 + r) * 2];
            fullMv[(i * 2 + r) * 2 + 1] = raster[(b * 2 + r) * 2 + 1];
            pipe_stage2(&c, raster, b, r, predMv + (i * 2 + r) * 2, intMv + (i * 2 + r) * 2, intCost + i * 2 + r,
                        qMv + (i * 2 + r) * 2, qCost + i * 2 + r);
        }
        refIdx[i] = pipe_final(&c, b, predMv + i * 4, qMv + i * 4, qCost + i * 2, candCost + (size_t)i * prm->ncand,
                               bestIdx + i, bestCost + i, bestMvx + i, bestMvy + i, predY, predU, predV, psy, psc);
    }
    free(have);
    free(raster);
}

/* ---------------------------------------------------------------------------
 * Quantisation and inverse quantisation of one transform block, coefficient
 * group by coefficient group: Xin266QuantInvQuant and Xin266QIQCodingGroup
 * (h266_encoder/source/h266_quant_inv_quant.c:27-143; slot pfXinQuantInvQuant,
 * typedef XinQuantInvQuant h266_encoder/include/h266_func_struct.h:181-196).
 * q = sign * ((|t| * qMult + qAdd) >> qShift), clipped to int16; r =
 * (q * iqMult + iqAdd) >> iqShift, clipped to int16 (MIN/MAX_COEFF_VALUE,
 * video_api/video_common/include/xin_typedef.h:80-81).  Bit k of *nzMap is set
 * when coefficient group k (raster order of the cgW x cgH groups) holds a
 * non-zero q: the reference walks the groups in reverse raster order and
 * shifts the map left before each, which lands group k on bit k.
 * ------------------------------------------------------------------------- */
void or_QuantInvQuant(int32_t *q, const int32_t *t, int32_t *r, intptr_t stride, uint32_t w, uint32_t h,
                      uint32_t cgw, uint32_t cgh, int32_t qMult, int32_t qAdd, int32_t qShift, int32_t iqMult,
                      int32_t iqAdd, int32_t iqShift, uint64_t *nzMap)
{
    const uint32_t ncx = w / cgw, ncy = h / cgh;
    uint64_t map = 0;
    for (uint32_t cy = 0; cy < ncy; cy++)
        for (uint32_t cx = 0; cx < ncx; cx++) {
            int nz = 0;
            for (uint32_t y = cy * cgh; y < (cy + 1) * cgh; y++)
                for (uint32_t x = cx * cgw; x < (cx + 1) * cgw; x++) {
                    const intptr_t i = (intptr_t)y * stride + x;
                    const int32_t tc = t[i], sign = tc < 0 ? -1 : 1, a = tc < 0 ? -tc : tc;
                    /* int32 arithmetic as the reference's SINT32 expressions */
                    int32_t qc = (int32_t)(((uint32_t)a * (uint32_t)qMult + (uint32_t)qAdd)) >> qShift;
                    qc *= sign;
                    qc = qc < -32768 ? -32768 : qc > 32767 ? 32767 : qc;
                    q[i] = qc;
                    int32_t rc = (int32_t)((uint32_t)qc * (uint32_t)iqMult + (uint32_t)iqAdd) >> iqShift;
                    rc = rc < -32768 ? -32768 : rc > 32767 ? 32767 : rc;
                    r[i] = rc;
                    nz |= qc != 0;
                }
            if (nz) map |= 1ull << (cy * ncx + cx);
        }
    *nzMap = map;
}

/* ----------------------------------------------------------------------------
 * Quantiser parameters of one TU: Xin266Transform
 * (h266_encoder/source/h266_trans_quant.c:104-120) with the quantParam table
 * of Xin266ConstructQuantParam (h266_encoder/source/h266_encoder_create.c:
 * 1677-1705; scales :94-104, the H.265/H.266 level scales and their
 * half-octave-shifted set for odd log2 areas), bit depth 10:
 *   adj = (lgW + lgH odd) and not skip, lgSize = (lgW + lgH) / 2,
 *   qShift = 14 + (15 - 10) + qp / 6 - lgSize - adj,
 *   iqShift = 6 - (15 - 10) + lgSize + adj,
 *   qAdd = rdoq and not skip ? 1 << (qShift - 1)
 *                            : (I frame ? 171 : 85) << (qShift - 9),
 *   iqAdd = not skip ? 1 << (iqShift - 1) : 1 << (iqShift - lgSize + 15 - 10 - 1).
 * ------------------------------------------------------------------------- */
static const int32_t kQScale[2][6] = {{26214, 23302, 20560, 18396, 16384, 14564},
                                      {18396, 16384, 14564, 13107, 11651, 10280}};
static const int32_t kIqScale[2][6] = {{40, 45, 51, 57, 64, 72}, {57, 64, 72, 80, 90, 102}};

void or_QuantParams(uint32_t desc, int32_t out[6])
{
    const int lgW = desc & 15, lgH = (desc >> 4) & 15, skip = ((desc >> 8) & 15) == 1;
    const int qp = (desc >> 18) & 63, rdoq = (desc >> 24) & 1, intra = (desc >> 25) & 1;
    const int lgTu = lgW + lgH, adj = (lgTu & 1) && !skip, lgSize = lgTu >> 1;
    const int per = qp / 6, rem = qp % 6;
    const int qShift = 14 + (15 - 10) + per - lgSize - adj;
    const int iqShift = 6 - (15 - 10) + lgSize + adj;
    out[0] = kQScale[adj][rem];
    out[1] = (rdoq && !skip) ? 1 << (qShift - 1) : (intra ? 171 : 85) << (qShift - 9);
    out[2] = qShift;
    out[3] = kIqScale[adj][rem] << per;
    out[4] = !skip ? 1 << (iqShift - 1) : 1 << ((iqShift - lgSize) + (15 - 10) - 1);
    out[5] = iqShift;
}

/* The batched quantiser (h266_trans_quant.c:204-222 per TU, the region being
 * XIN_MIN(width, 32) x XIN_MIN(height, 32); the reference zeroes the
 * dequantised coefficients past it at :163-188). */
void or_QuantBatch(const int32_t *t, const int64_t *offset, const uint32_t *desc, int32_t i0, int32_t i1,
                   int16_t *q, int16_t *r, uint64_t *nzMap)
{
    for (int32_t i = i0; i < i1; i++) {
        const uint32_t d = desc[i];
        const int w = 1 << (d & 15), h = 1 << ((d >> 4) & 15);
        const int rw = w < 32 ? w : 32, rh = h < 32 ? h : 32;
        int32_t p[6];
        or_QuantParams(d, p);
        int32_t *q32 = (int32_t *)malloc(sizeof(int32_t) * (size_t)w * h);
        int32_t *r32 = (int32_t *)malloc(sizeof(int32_t) * (size_t)w * h);
        memset(q32, 0, sizeof(int32_t) * (size_t)w * h);
        memset(r32, 0, sizeof(int32_t) * (size_t)w * h);
        or_QuantInvQuant(q32, t + offset[i], r32, w, (uint32_t)rw, (uint32_t)rh, 1u << ((d >> 12) & 7),
                         1u << ((d >> 15) & 7), p[0], p[1], p[2], p[3], p[4], p[5], nzMap + i);
        for (int k = 0; k < w * h; k++) {
            q[offset[i] + k] = (int16_t)q32[k];
            r[offset[i] + k] = (int16_t)r32[k];
        }
        free(q32);
        free(r32);
    }
}

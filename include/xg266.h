/*
 * xg266.h -- C-ABI of the MI355X (gfx950) VVC block-matching library.
 *
 * Two layers:
 *
 *  1. Per-call primitives with exactly the signatures of xin26x's pixel
 *     primitive table `xin_func_struct` (h266_encoder/include/h266_func_struct.h:719-799)
 *     and of its motion-estimation entry points.  Pointers are HOST pointers
 *     into the caller's buffers, as in the reference; each call stages the
 *     bytes it needs to HBM, runs one HIP kernel and returns synchronously.
 *     A C encoder assigns these to its table after Xin266FuncInit; the
 *     Xg266FuncInit helper fills the mirror table below.
 *
 *  2. Batched entry points over DEVICE pointers: all blocks of a frame (or of
 *     many frames of many streams) in one persistent kernel per stage, fed by
 *     per-shape task lists, optionally replayed from a captured HIP graph.
 *
 * Types: PIXEL is uint16 (10-bit samples in [0,1023]), COEFF is int32; these
 * match the reference's ENABLE_10BIT_ENCODER build
 * (video_api/video_common/include/xin_typedef.h:72-94).  Strides are in
 * elements.  All functions are thread-safe (one HIP stream per calling host
 * thread).  Per-call primitives return void like the reference; a failure
 * sets the value returned by xg_last_error() and reports through the logger.
 */
#ifndef XG266_H
#define XG266_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint16_t xg_pixel;
typedef int32_t  xg_coeff;

/* ------------------------------------------------------------------------
 * Status, errors, logging
 * ---------------------------------------------------------------------- */
typedef enum xg_status {
    XG_OK = 0,
    XG_ERR_ARG = 1,        /* invalid argument (size, pointer, range)  */
    XG_ERR_HIP = 2,        /* a HIP runtime call failed                 */
    XG_ERR_NOMEM = 3,      /* device or pinned allocation failed        */
    XG_ERR_BUDGET = 4,     /* admission refused: HBM budget exhausted   */
    XG_ERR_STATE = 5       /* call not valid in the object's state      */
} xg_status;

/* Last error of the calling thread (XG_OK if none since xg_clear_error). */
int32_t xg_last_error(void);
void    xg_clear_error(void);
/* Logger callback: level 0 error, 1 warning, 2 info.  NULL restores stderr. */
typedef void (*xg_log_fn)(int32_t level, const char *msg);
void    xg_set_logger(xg_log_fn fn);
/* Library version string and the device architecture it was built for. */
const char *xg_version(void);

/* ------------------------------------------------------------------------
 * 1. Per-call primitives (host pointers, synchronous)
 * ---------------------------------------------------------------------- */

/* XinComputeDist (h266_func_struct.h:56-63).  pfXinComputeSad[] slot;
 * replaces the closed Xin26xComputeSad (h26x_core/include/h26x_compute_dist.h:27). */
void Xg266ComputeSad(xg_pixel *input, intptr_t inputStride, xg_pixel *pred, intptr_t predStride,
                     uint32_t width, uint32_t height, uint32_t *sad);
/* pfXinComputeSatd[] slot; replaces the closed Xin26xComputeSatd
 * (h26x_compute_dist.h:806-937).  Hadamard 8x8 tiles, 4x4 when a side is 4. */
void Xg266ComputeSatd(xg_pixel *input, intptr_t inputStride, xg_pixel *pred, intptr_t predStride,
                      uint32_t width, uint32_t height, uint32_t *satd);
/* XinComputeSadx8 (h266_func_struct.h:87-94); replaces Xin26xComputeSadx8
 * (h26x_compute_dist.h:400-735). */
void Xg266ComputeSadx8(xg_pixel *input, intptr_t inputStride, xg_pixel *ref[8], intptr_t refStride,
                       uint32_t width, uint32_t height, uint32_t *sad);
/* XinComputeSadx5 (h266_func_struct.h:106-117). */
void Xg266ComputeSadx5(xg_pixel *input, intptr_t inputStride, xg_pixel *ref0, xg_pixel *ref1,
                       xg_pixel *ref2, xg_pixel *ref3, xg_pixel *ref4, intptr_t refStride,
                       uint32_t width, uint32_t height, uint32_t *sad);
/* XinComputeSadx3 (h266_func_struct.h:119-128). */
void Xg266ComputeSadx3(xg_pixel *input, intptr_t inputStride, xg_pixel *ref0, xg_pixel *ref1,
                       xg_pixel *ref2, intptr_t refStride, uint32_t width, uint32_t height,
                       uint32_t *sad);
/* XinComputeAvgDist (h266_func_struct.h:65-73): pfXinComputeAvgSad[] and
 * pfXinComputeAvgSatd[] slots, replacing the closed Xin26xComputeAvgSad /
 * Xin26xComputeAvgSatd (h26x_compute_dist.h:153-243, 824-900).  SAD / SATD
 * of the input against the rounded average (a + b + 1) >> 1 of two
 * predictions (the bi-prediction cost of h266_analyze_inter_cu.c:720, 817, 1214). */
void Xg266ComputeAvgSad(xg_pixel *input, intptr_t inputStride, xg_pixel *refA, xg_pixel *refB,
                        intptr_t refStride, uint32_t width, uint32_t height, uint32_t *sad);
void Xg266ComputeAvgSatd(xg_pixel *input, intptr_t inputStride, xg_pixel *refA, xg_pixel *refB,
                         intptr_t refStride, uint32_t width, uint32_t height, uint32_t *satd);
/* XinComputeAvgSadx8 (h266_func_struct.h:96-104); replaces Xin26xComputeAvgSadx8
 * (h26x_compute_dist.h:735-795): the averaged SAD for 8 prediction pairs. */
void Xg266ComputeAvgSadx8(xg_pixel *input, intptr_t inputStride, xg_pixel *refA[8], xg_pixel *refB[8],
                          intptr_t refStride, uint32_t width, uint32_t height, uint32_t *sad);
/* XinQuantInvQuant (h266_func_struct.h:181-196): pfXinQuantInvQuant[] slot,
 * replacing Xin266QuantInvQuant (h266_encoder/source/h266_quant_inv_quant.c:80-143).
 * Per coefficient group of cgWidth x cgHeight (at most 64 groups):
 * q = sign(t) * ((|t| * qMult + qAdd) >> qShift) and r = (q * iqMult + iqAdd) >> iqShift,
 * both clipped to int16; bit k of *nzCGBitMapRs is set when group k (raster
 * order) holds a non-zero q. */
void Xg266QuantInvQuant(xg_coeff *qCoeff, xg_coeff *tCoeff, xg_coeff *rCoeff, intptr_t coeffStride,
                        uint32_t width, uint32_t height, uint32_t cgWidth, uint32_t cgHeight, int32_t qMult,
                        int32_t qAdd, int32_t qShift, int32_t iqMult, int32_t iqAdd, int32_t iqShift,
                        uint64_t *nzCGBitMapRs);
/* XinComputeHada8x8 (h266_func_struct.h:448-451); replaces
 * Xin26xComputeHada8x8 (h26x_compute_dist.h:937). */
void Xg266ComputeHada8x8(xg_pixel *input, intptr_t inputStride, uint32_t *satd);

/* XinMotionInterp (h266_func_struct.h:130-138): luma slots of
 * pfXinLumaInterp[!!fx][!!fy][] (h266_func_init.c:424-458), replacing
 * h266_core/source/h266_luma_mc.c:72-428. */
void Xg266InterpCopy(const xg_pixel *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                     int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266LumaInterpHor(const xg_pixel *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                        int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266LumaInterpVet(const xg_pixel *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                        int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266LumaInterpHorVet(const xg_pixel *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                           int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
/* XinMotionInterpS16 (h266_func_struct.h:140-148): pfXinLumaInterpS16[][][]
 * (h266_func_init.c:460-494), replacing h266_luma_mc.c:216-273, 430-559. */
void Xg266InterpCopyU8S16(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride,
                          int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266LumaInterpHorU8S16(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride,
                             int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266LumaInterpVetU8S16(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride,
                             int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266LumaInterpHorVetU8S16(const xg_pixel *src, intptr_t srcStride, int16_t *dst,
                                intptr_t dstStride, int32_t frac, uint32_t width, uint32_t height,
                                uint32_t filterIndex);
/* Second-pass helpers with S16 input (h266_luma_mc.c:275-392). */
void Xg266LumaInterpVetS16U8(const int16_t *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                             int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266LumaInterpVetS16S16(const int16_t *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride,
                              int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
/* Chroma: pfXinChromaInterp / pfXinChromaInterpS16 (h266_func_struct.h:741,744),
 * replacing h266_core/source/h266_chroma_mc.c:62-384 (frac = (fy<<5)|fx). */
void Xg266ChromaInterpHor(const xg_pixel *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                          int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266ChromaInterpVet(const xg_pixel *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                          int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266ChromaInterpHorVet(const xg_pixel *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                             int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266ChromaInterpHorU8S16(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride,
                               int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266ChromaInterpVetU8S16(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride,
                               int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266ChromaInterpHorVetU8S16(const xg_pixel *src, intptr_t srcStride, int16_t *dst,
                                  intptr_t dstStride, int32_t frac, uint32_t width, uint32_t height,
                                  uint32_t filterIndex);
void Xg266ChromaInterpVetS16U8(const int16_t *src, intptr_t srcStride, xg_pixel *dst, intptr_t dstStride,
                               int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
void Xg266ChromaInterpVetS16S16(const int16_t *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride,
                                int32_t frac, uint32_t width, uint32_t height, uint32_t filterIndex);
/* XinMotionInterpAvg / XinMotionInterpWeight (h266_func_struct.h:150-173):
 * pfXinInterpAvg / pfXinInterpWeight, replacing h266_luma_mc.c:561-629. */
void Xg266InterpAvgS16U8(const int16_t *src0, intptr_t src0Stride, const int16_t *src1,
                         intptr_t src1Stride, xg_pixel *dst, intptr_t dstStride, uint32_t width,
                         uint32_t height);
void Xg266InterpWeightS16U8(const int16_t *src0, intptr_t src0Stride, const int16_t *src1,
                            intptr_t src1Stride, xg_pixel *dst, intptr_t dstStride, uint32_t width,
                            uint32_t height, int32_t weightA, int32_t weightB);

/* XinFor1dDct (h266_func_struct.h:198-204): pfXinFor1dDct[type][lgN]
 * (h266_func_init.c:349-374), replacing the closed Xin26xFDct2P2..P64,
 * Xin26xFDst7P4..P32, Xin26xFDct8P4..P32 (h26x_core/include/h26x_forward_1d_trans.h:31-141).
 *   out[k*outputStride + j] = (sum_n T[k][n]*in[j*inputStride + n] + 2^(shift-1)) >> shift,
 *   j < line; outputs past the VVC zero-out (32 for DCT-II-64, 16 for 32-point
 *   DST-VII/DCT-VIII) are written as 0. */
void Xg26xFDct2P2(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct2P4(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct2P8(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct2P16(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct2P32(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct2P64(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDst7P4(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDst7P8(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDst7P16(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDst7P32(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct8P4(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct8P8(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct8P16(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
void Xg26xFDct8P32(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t shift);
/* XinFor2dDct2 (h266_func_struct.h:206-215): pfXinFor2dDct2[lgW][lgH]
 * (h266_func_init.c:72-85), replacing the closed Xin26xFDct2WxH
 * (h26x_core/include/h26x_forward_2d_trans.h:23).  tempBuf is unused. */
void Xg26xFDct2WxH(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                   uint32_t width, uint32_t height, xg_coeff *tempBuf, int32_t shift1st,
                   int32_t shift2nd);
/* XinBlockSubForDct: pfXinBlockSubForDct[] (h266_sub_dct.c:74-82), replacing
 * the closed Xin26xBlockSubForDct (h26x_core/include/h26x_block_utility.h:151). */
void Xg26xBlockSubForDct(xg_pixel *input, intptr_t inputStride, xg_pixel *recon, intptr_t reconStride,
                         xg_coeff *diff, intptr_t diffStride, uint32_t width, uint32_t height);
/* Residual + transform of one TU with the dispatch of Xin266FDctWxH
 * (h266_encoder/source/h266_forward_trans.c:45-152): mtsIdx per
 * h26x_trans_context.h:42-51, implicit MTS for intra luma, shifts
 * lgW+bitDepth-9 and lgH+6.  Transform skip (mtsIdx 1) returns XG_ERR_ARG. */
int32_t Xg266FDctWxH(const xg_coeff *input, intptr_t inputStride, xg_coeff *output,
                     intptr_t outputStride, uint32_t width, uint32_t height, uint32_t mtsIdx,
                     int32_t isIntraLuma, int32_t enableMts);

/* XinInv1dDct (h266_func_struct.h:280-287): pfXinInv1dDct[type][lgN]
 * (h266_func_init.c:376-401), replacing the closed Xin26xIDct2P2..P64,
 * Xin26xIDst7P4..P32, Xin26xIDct8P4..P32 (h26x_core/include/h26x_inverse_1d_trans.h:23-212).
 *   out[j*outputStride + i] = Clip3(-32768, 32767,
 *       (sum_{k<nz} T[k][i]*in[k*inputStride + j] + 2^(shift-1)) >> shift)
 * for j < line - skipLine; 0 for the last skipLine lines, whose input is not
 * read.  nz is the H.266 zero-out bound (32 for DCT-II-64, 16 for 32-point
 * DST-VII/DCT-VIII): coefficients past it are not read. */
void Xg26xIDct2P2(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct2P4(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct2P8(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct2P16(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct2P32(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct2P64(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDst7P4(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDst7P8(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDst7P16(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDst7P32(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct8P4(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct8P8(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct8P16(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
void Xg26xIDct8P32(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride, int32_t line, int32_t skipLine, int32_t shift);
/* XinInv2dDct2 (h266_func_struct.h:269-278): pfXinInv2dDct2[lgW][lgH]
 * (h266_func_init.c:72-85), replacing Xin266IDct2WxH
 * (h266_encoder/source/h266_inverse_trans.c:51-89): vertical then horizontal
 * DCT-II, coefficient columns past 32 skipped.  tempBuf is unused. */
void Xg266IDct2WxH(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                   uint32_t width, uint32_t height, xg_coeff *tempBuf, int32_t shift1st, int32_t shift2nd);
/* XinInv2dSkip (h266_func_struct.h:635-641): pfXinInv2dSkip[] (h266_func_init.c:410-419),
 * replacing Xin266ISkipWxH (h266_inverse_trans.c:91-113):
 * out = (int16_t)(in >> (7 - ((lgW + lgH) >> 1))), sizes 1..128. */
void Xg266ISkipWxH(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                   uint32_t width, uint32_t height);
/* Inverse transform of one TU with the dispatch of Xin266IDctWxH
 * (h266_encoder/source/h266_inverse_trans.c:115-229): mtsIdx as
 * Xg266FDctWxH, implicit intra-luma DST-VII, shifts 7 and 20 - 10, zero-out
 * skip of 16 lines for a 32-wide DST-VII/DCT-VIII, transform skip for
 * mtsIdx 1.  Widths or heights of 1 outside transform skip return XG_ERR_ARG
 * (the reference leaves that result in its section scratch buffer). */
int32_t Xg266IDctWxH(const xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                     uint32_t width, uint32_t height, uint32_t mtsIdx, int32_t isIntraLuma, int32_t enableMts);
/* XinBlockAddForDct (h266_func_struct.h:530-538): pfXinBlockAddForDct[]
 * (h266_func_init.c:161-166), replacing the closed Xin26xBlockRecon:
 * recon = Clip3(0, 1023, pred + residual). */
void Xg26xBlockRecon(xg_coeff *input, intptr_t inputStride, xg_pixel *pred, intptr_t predStride,
                     xg_pixel *recon, intptr_t reconStride, uint32_t width, uint32_t height);
/* Xin266IDctAdd (h266_encoder/source/h266_idct_add.c:45-104): inverse
 * transform then reconstruction, sizes up to 64x64.  The LMCS chroma
 * residual scaling the reference applies when LMCS is enabled
 * (:84-92) is not part of this entry. */
int32_t Xg266IDctAdd(const xg_coeff *input, intptr_t inputStride, const xg_pixel *pred, intptr_t predStride,
                     xg_pixel *output, intptr_t outputStride, uint32_t width, uint32_t height, uint32_t mtsIdx,
                     int32_t isIntraLuma, int32_t enableMts);

/* XinBiliInterp (h266_func_struct.h:150-157): pfXinBiliInterp[!!fracX][!!fracY][lgW]
 * (h266_func_init.c:496-530), replacing Xin266BiliInterpCopy / Hor / Vet / HorVet
 * (h266_core/source/h266_bilinear_filter.c:46-260): the DMVR bilinear
 * prediction, width + 4 columns of int16 per row; frac = fracY << 4 | fracX.
 * 10-bit: Copy keeps the samples, each filter pass is ((16-f)*a + f*b + 8) >> 4.
 * Widths 1..128, heights 1..132. */
void Xg266BiliInterpCopy(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride, int32_t frac,
                         uint32_t width, uint32_t height);
void Xg266BiliInterpHor(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride, int32_t frac,
                        uint32_t width, uint32_t height);
void Xg266BiliInterpVet(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride, int32_t frac,
                        uint32_t width, uint32_t height);
void Xg266BiliInterpHorVet(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride, int32_t frac,
                           uint32_t width, uint32_t height);
/* XinComputeDistS16 (h266_func_struct.h:540-548): pfXinComputeSadS16[] (h266_func_init.c:118-125),
 * replacing the closed Xin26xComputeSadS16 (h26x_core/include/h26x_compute_dist.h:942-950):
 * SAD of rows 0, 2^subShift, ... scaled back by << subShift. */
void Xg26xComputeSadS16(int16_t *input, intptr_t inputStride, int16_t *pred, intptr_t predStride, uint32_t width,
                        uint32_t height, uint32_t subShift, uint32_t *sad);
/* One bi-predicted block for decoder-side MV refinement: the fields
 * Xin266DeriveMvFromDmvr (h266_encoder/source/h266_inter_prediction.c:976-1195)
 * reads from its section, CU and PU. */
typedef struct xg_dmvr {
    const xg_pixel *ref0; intptr_t stride0;     /* sample (0, 0) of each padded reference plane (luma) */
    const xg_pixel *ref1; intptr_t stride1;
    int32_t  cuPelX, cuPelY;                    /* block position, luma samples */
    uint32_t width, height;                     /* 8..128, powers of two */
    int32_t  mv0x, mv0y, mv1x, mv1y;            /* list 0 / list 1 MVs, 1/16 pel */
    int16_t  minMvX, minMvY, maxMvX, maxMvY;    /* integer-pel MV window (secSet->minMv / maxMv) */
} xg_dmvr;
/* Replaces Xin266DeriveMvFromDmvr: one (x, y) delta in 1/16 pel per subblock
 * of min(16, side) in raster order (mvdL0SubPu), host pointers. */
int32_t Xg266DeriveMvFromDmvr(const xg_dmvr *dmvr, int16_t *mvdL0SubPu);

/* Mirror of the hot-path slots of xin_func_struct (h266_func_struct.h:719-799),
 * same names, same index conventions (index = log2 of the width). */
typedef void (*XgComputeDist)(xg_pixel *, intptr_t, xg_pixel *, intptr_t, uint32_t, uint32_t, uint32_t *);
typedef void (*XgComputeAvgDist)(xg_pixel *, intptr_t, xg_pixel *, xg_pixel *, intptr_t, uint32_t, uint32_t,
                                 uint32_t *);
typedef void (*XgComputeSadx8)(xg_pixel *, intptr_t, xg_pixel *[8], intptr_t, uint32_t, uint32_t, uint32_t *);
typedef void (*XgComputeAvgSadx8)(xg_pixel *, intptr_t, xg_pixel *[8], xg_pixel *[8], intptr_t, uint32_t,
                                  uint32_t, uint32_t *);
typedef void (*XgComputeSadx5)(xg_pixel *, intptr_t, xg_pixel *, xg_pixel *, xg_pixel *, xg_pixel *,
                               xg_pixel *, intptr_t, uint32_t, uint32_t, uint32_t *);
typedef void (*XgComputeSadx3)(xg_pixel *, intptr_t, xg_pixel *, xg_pixel *, xg_pixel *, intptr_t,
                               uint32_t, uint32_t, uint32_t *);
typedef void (*XgMotionInterp)(const xg_pixel *, intptr_t, xg_pixel *, intptr_t, int32_t, uint32_t,
                               uint32_t, uint32_t);
typedef void (*XgMotionInterpS16)(const xg_pixel *, intptr_t, int16_t *, intptr_t, int32_t, uint32_t,
                                  uint32_t, uint32_t);
typedef void (*XgMotionInterpAvg)(const int16_t *, intptr_t, const int16_t *, intptr_t, xg_pixel *,
                                  intptr_t, uint32_t, uint32_t);
typedef void (*XgMotionInterpWeight)(const int16_t *, intptr_t, const int16_t *, intptr_t, xg_pixel *,
                                     intptr_t, uint32_t, uint32_t, int32_t, int32_t);
typedef void (*XgFor1dDct)(xg_coeff *, intptr_t, xg_coeff *, intptr_t, int32_t, int32_t);
typedef void (*XgFor2dDct2)(xg_coeff *, intptr_t, xg_coeff *, intptr_t, uint32_t, uint32_t, xg_coeff *,
                            int32_t, int32_t);
typedef void (*XgBlockSubForDct)(xg_pixel *, intptr_t, xg_pixel *, intptr_t, xg_coeff *, intptr_t,
                                 uint32_t, uint32_t);
typedef void (*XgInv1dDct)(xg_coeff *, intptr_t, xg_coeff *, intptr_t, int32_t, int32_t, int32_t);
typedef void (*XgInv2dDct2)(xg_coeff *, intptr_t, xg_coeff *, intptr_t, uint32_t, uint32_t, xg_coeff *,
                            int32_t, int32_t);
typedef void (*XgInv2dSkip)(xg_coeff *, intptr_t, xg_coeff *, intptr_t, uint32_t, uint32_t);
typedef void (*XgBlockAddForDct)(xg_coeff *, intptr_t, xg_pixel *, intptr_t, xg_pixel *, intptr_t, uint32_t,
                                 uint32_t);
typedef void (*XgBiliInterp)(const xg_pixel *, intptr_t, int16_t *, intptr_t, int32_t, uint32_t, uint32_t);
typedef void (*XgComputeDistS16)(int16_t *, intptr_t, int16_t *, intptr_t, uint32_t, uint32_t, uint32_t,
                                 uint32_t *);
typedef void (*XgComputeHada8x8)(xg_pixel *, intptr_t, uint32_t *);
typedef void (*XgQuantInvQuant)(xg_coeff *, xg_coeff *, xg_coeff *, intptr_t, uint32_t, uint32_t, uint32_t, uint32_t,
                                int32_t, int32_t, int32_t, int32_t, int32_t, int32_t, uint64_t *);

#define XG_BLOCK_NUM 8
#define XG_TRANS_TYPE 3        /* XIN_DCT2 = 0, XIN_DCT8 = 1, XIN_DST7 = 2 */

typedef struct xg_func_struct {
    XgComputeDist        pfXinComputeSad[XG_BLOCK_NUM];
    XgComputeDist        pfXinComputeSatd[XG_BLOCK_NUM];
    XgComputeDist        pfXinComputeDist[XG_BLOCK_NUM];
    XgComputeAvgDist     pfXinComputeAvgSad[XG_BLOCK_NUM];
    XgComputeAvgDist     pfXinComputeAvgSatd[XG_BLOCK_NUM];
    XgComputeAvgDist     pfXinComputeAvgDist[XG_BLOCK_NUM];
    XgComputeSadx8       pfXinComputeSadx8[XG_BLOCK_NUM];
    XgComputeAvgSadx8    pfXinComputeAvgSadx8[XG_BLOCK_NUM];
    XgComputeSadx5       pfXinComputeSadx5[XG_BLOCK_NUM];
    XgComputeSadx3       pfXinComputeSadx3[XG_BLOCK_NUM];
    XgMotionInterp       pfXinLumaInterp[2][2][XG_BLOCK_NUM];
    XgMotionInterp       pfXinChromaInterp[2][2][XG_BLOCK_NUM];
    XgMotionInterpS16    pfXinLumaInterpS16[2][2][XG_BLOCK_NUM];
    XgMotionInterpS16    pfXinChromaInterpS16[2][2][XG_BLOCK_NUM];
    XgMotionInterpAvg    pfXinInterpAvg[XG_BLOCK_NUM];
    XgMotionInterpWeight pfXinInterpWeight[XG_BLOCK_NUM];
    XgFor2dDct2          pfXinFor2dDct2[XG_BLOCK_NUM][XG_BLOCK_NUM];
    XgFor1dDct           pfXinFor1dDct[XG_TRANS_TYPE][XG_BLOCK_NUM];
    XgBlockSubForDct     pfXinBlockSubForDct[XG_BLOCK_NUM];
    XgQuantInvQuant      pfXinQuantInvQuant[XG_BLOCK_NUM];
    XgInv2dSkip          pfXinInv2dSkip[XG_BLOCK_NUM];
    XgInv2dDct2          pfXinInv2dDct2[XG_BLOCK_NUM][XG_BLOCK_NUM];
    XgInv1dDct           pfXinInv1dDct[XG_TRANS_TYPE][XG_BLOCK_NUM];
    XgBlockAddForDct     pfXinBlockAddForDct[XG_BLOCK_NUM];
    XgBiliInterp         pfXinBiliInterp[2][2][XG_BLOCK_NUM];
    XgComputeDistS16     pfXinComputeSadS16[XG_BLOCK_NUM];
    XgComputeHada8x8     pfXinComputeHada8x8;
} xg_func_struct;

/* Fills every slot above with this library's functions (NULL where the
 * reference table holds NULL: SATD and averaged SAD/SATD widths 1-2,
 * quantisation width 128, reconstruction widths 1 and 128,
 * DST-VII/DCT-VIII outside 4..32, DCT-II 1 and 128), then binds
 * pfXinComputeDist / pfXinComputeAvgDist to SAD or SATD as Xin266FuncUpdate
 * does (h266_encoder_create.c:3332-3400).  Initialises the
 * HIP device on the first call. */
int32_t Xg266FuncInit(xg_func_struct *funcSet, int32_t satdMd);

/* Motion-estimation boundary.  Field-for-field the inputs and outputs the
 * reference's closed Xin26xMotionEst* read and write in xin_me_struct
 * (h26x_core/include/h26x_me_struct.h:60-131); MVs: minMv/maxMv/meCand in
 * integer pel, predMv/bestMv in quarter pel.  `ref` points at the co-located
 * block in a padded reference picture (the caller guarantees padding for
 * [minMv, maxMv] plus 4 samples of filter reach). */
typedef struct xg_mv { int16_t mvX, mvY; } xg_mv;
typedef struct xg_me_struct {
    uint32_t  width, height;
    xg_mv     minMv, maxMv;
    xg_mv     meCand[12];
    uint32_t  candNum;
    xg_mv     bestMv, predMv;
    uint32_t  bestCost, bestSad;
    int32_t   searchWidth, searchHeight;
    uint32_t  sadLambda;
    xg_pixel *input;  intptr_t inputStride;
    xg_pixel *ref;    intptr_t refStride;
} xg_me_struct;

/* Replaces Xin26xMotionEstFull (h26x_core/include/h26x_motion_est.h:26):
 * exhaustive raster search of +/-searchWidth x +/-searchHeight around the
 * rounded predictor, clipped to [minMv, maxMv]; cost (SAD<<8) + bits*lambda. */
void Xg26xMotionEstFull(xg_me_struct *meSet);
/* Replaces Xin26xMotionEstDia (h26x_motion_est.h:35): start candidates then
 * small-diamond descent. */
void Xg26xMotionEstDia(xg_me_struct *meSet);
/* Replaces Xin266FullSubPelMe (h266_encoder/source/h266_sub_pel_refine.c:81-212):
 * half-pel then quarter-pel ring with 8-tap interpolation and SAD. */
void Xg266FullSubPelMe(xg_me_struct *meSet);

/* ------------------------------------------------------------------------
 * 2. Batched entry points (device pointers)
 * ---------------------------------------------------------------------- */

/* One plane: `data` points at sample (0,0) of a buffer padded by at least
 * `pad` samples on every side (edge-replicated by the caller). */
typedef struct xg_plane {
    xg_pixel *data;
    intptr_t  stride;
    int32_t   width, height, pad;
} xg_plane;

/* 4:2:0 picture: plane[0] luma, plane[1] Cb, plane[2] Cr. */
typedef struct xg_picture { xg_plane plane[3]; } xg_picture;

/* One coding block: luma position/size and its motion (1/16 pel). */
typedef struct xg_block {
    int16_t x, y;          /* luma position                                   */
    uint8_t w, h;          /* 4..64, powers of two                            */
    uint8_t pic, ref;      /* picture-set slot of the current / reference pic */
    int32_t mvx, mvy;      /* base MV of the candidate pattern, 1/16 pel      */
    int32_t mvpx, mvpy;    /* MV predictor for the rate term, 1/16 pel        */
} xg_block;

/* Fractional-refinement configuration (per call). */
#define XG_MAX_CAND 64
typedef struct xg_refine_cfg {
    int32_t  ncand;                  /* candidates per block, <= XG_MAX_CAND    */
    int16_t  offx[XG_MAX_CAND];      /* candidate offsets around the base MV, */
    int16_t  offy[XG_MAX_CAND];      /*   1/16 pel, |off| <= 64               */
    uint32_t lambda;                 /* sadLambda (h266_enc_init.c:86-105)    */
    int32_t  ctuSize;                /* MV clip window (h266_enc_init.c:1462) */
    int32_t  useChroma;              /* add Cb+Cr SATD into each cost          */
} xg_refine_cfg;

/* Result of one block. */
typedef struct xg_block_result {
    int32_t  mvx, mvy;     /* best MV after clipping, 1/16 pel */
    uint32_t cost;         /* (SATD << 8) + bits * lambda      */
    int32_t  cand;         /* index of the best candidate      */
} xg_block_result;

typedef enum xg_mode {
    XG_MODE_PLAIN = 0,     /* one kernel launch per block shape                  */
    XG_MODE_BATCHED = 1,   /* one persistent kernel per stage over all shapes    */
    XG_MODE_GRAPH = 2      /* the batched stages replayed from a captured graph  */
} xg_mode;

typedef struct xg_session xg_session;

/* A session owns a HIP stream, the per-shape task lists and the result
 * buffers for up to maxBlocks blocks.  Pictures are passed per call through a
 * picture set of up to 128 slots (xg_block.pic / .ref index it). */
int32_t xg_session_create(xg_session **out, int32_t maxBlocks);
int32_t xg_session_destroy(xg_session *s);
/* HIP stream of the session (hipStream_t) so callers can order their work. */
void   *xg_session_stream(xg_session *s);
/* Installs a picture in slot `slot` (0..127).  Device pointers. */
int32_t xg_session_set_picture(xg_session *s, int32_t slot, const xg_picture *pic);
/* Uploads the block list (host memory) and builds per-shape task lists:
 * tasks sorted by shape, largest first, each covering `candPerTask`
 * candidates of one block (per-shape tile table, see xg_session_set_tile). */
int32_t xg_session_set_blocks(xg_session *s, const xg_block *blocks, int32_t n);
/* Per-shape tile choice: candidates per task for shape (lgW-2, lgH-2). */
int32_t xg_session_set_tile(xg_session *s, int32_t lgW, int32_t lgH, int32_t candPerTask);
int32_t xg_session_get_tile(xg_session *s, int32_t lgW, int32_t lgH);
/* Scheduling of the batched form: persistent workgroups per CU, and
 * `regions` (1..8) horizontal bands of equal estimated cost, band r run by the
 * workgroups g with g % regions == r (workgroup g runs on XCD g % 8).  Tasks
 * are assigned to workgroups on the host (heaviest first, least loaded
 * workgroup), so the launch needs no work-queue atomics. */
int32_t xg_session_set_sched(xg_session *s, int32_t regions, int32_t wgPerCU);
/* Tunes the batched form on the installed blocks and pictures: first the
 * scheduling (spatial queues x workgroups per CU), then candidates per task
 * for each block shape present by coordinate descent, heaviest shape first;
 * every choice is timed on whole-frame launches (costs only) with HIP events.
 * Writes the schedule and tile table to `path` when not NULL
 * (xg_session_load_tiles reads it back).  Results never depend on the tiles. */
int32_t xg_session_autotune(xg_session *s, const xg_refine_cfg *cfg, int32_t reps, const char *path);
int32_t xg_session_load_tiles(xg_session *s, const char *path);

/* Prediction target of current-picture slot `slot` (device planes of the
 * picture's size, unpadded).  Used by xg_refine_frac when its `pred` is NULL. */
int32_t xg_session_set_pred(xg_session *s, int32_t slot, const xg_picture *pred);
/* Integer full search of every installed block (one workgroup per block):
 * +/-range around the centre ((mvp >> 2) + 2) >> 2 (the predictor in quarter
 * pel rounded to integer pel as h266_analyze_inter_cu.c:271-272, clamped into
 * the window), clipped to the CU MV window (h266_enc_init.c:1462-1465), cost
 * (SAD << 8) + (EG(4*dx - px) + EG(4*dy - py)) * sadLambda with the predictor
 * in quarter pel (mvp >> 2), raster order, earliest wins ties.  Writes each
 * block's base MV (mvx, mvy, 1/16 pel) in the session's device block list,
 * where xg_refine_frac then starts.  Enqueued on the session stream.  Block
 * widths are powers of two 4..64 at even x in dword-aligned planes (the
 * current block is read as sample pairs); range <= 32. */
int32_t xg_search_blocks(xg_session *s, int32_t range, uint32_t sadLambda, int32_t ctuSize);
/* Fractional refinement (per block: ncand candidates, luma 8-tap + chroma
 * 4-tap interpolation, SATD, cost) followed by motion compensation of the best
 * candidate into the prediction picture: `pred` for every block when not NULL,
 * else each block's slot target set by xg_session_set_pred.  Enqueued on the
 * session stream; returns immediately. */
int32_t xg_refine_frac(xg_session *s, const xg_refine_cfg *cfg, const xg_picture *pred,
                       int32_t mode);
/* Device pointers of the last results: candCost[block*ncand + c], best[block]. */
int32_t xg_session_outputs(xg_session *s, uint32_t **candCost, xg_block_result **best);
int32_t xg_session_sync(xg_session *s);
/* Waits for the session stream and copies the last results to host memory
 * (either pointer may be NULL): candCost[nblocks*ncand], best[nblocks]. */
int32_t xg_session_read(xg_session *s, uint32_t *candCost, xg_block_result *best);
/* Average device time (ms) per launch since the last reset, measured with HIP
 * events on the session stream while timing is enabled (xg_session_set_timing):
 * stage 0 the refinement kernel(s) of the eager modes, 1 a whole graph replay,
 * 2 the forward-transform kernel (xg_fdct_run), 3 the integer-search kernel
 * (xg_search_blocks, xg_full_search_frame, the full search of xg_me_pipeline),
 * 4 a whole xg_me_pipeline frame, 5 the reconstruction kernel (xg_recon_run),
 * 6 the DMVR kernel (xg_dmvr_run), 7 the bi-prediction distortion kernel
 * (xg_bipred_run), 8 the
 * quantisation kernel (xg_quant_run), 9 a whole xg_mctf_me frame, 10 its
 * finest level (8x8 blocks with the fractional rounds). */
int32_t xg_session_kernel_time(xg_session *s, int32_t stage, float *avgMs, int32_t *count);
int32_t xg_session_reset_timers(xg_session *s);
int32_t xg_session_set_timing(xg_session *s, int32_t enable);

/* ------------------------------------------------------------------------
 * 3. Serving: continuous batching of frames from many concurrent streams
 * ---------------------------------------------------------------------- */

/* A server keeps each admitted stream's state resident in HBM: its reference
 * and current pictures, its prediction picture, its block partition with the
 * persistent kernel's task assignment, and an MV field of 4x4 units holding
 * the final MVs of its last frame.  A dispatcher thread, whenever one of two
 * device sessions is free, batches the oldest pending frame of every stream
 * without a frame in flight into shared launches: one padding launch, one
 * integer-search launch and one refinement+MC launch over all their blocks.
 * Frames that arrive while a batch is on the GPU join the next batch; the two
 * sessions alternate so consecutive batches overlap on the device.  A
 * stream's first frame uses itself as reference; afterwards each processed
 * frame becomes the reference of the next.
 *
 * Predictors are resident too: every processed frame writes each block's
 * final MV (1/16 pel) into all its 4x4 units of the stream's MV field (zero
 * before the first frame).  The next frame's blocks take their AMVP candidates
 * from that field as xg_amvp_from_field does, and the predictor (the centre of
 * the integer search and the MVD origin of every cost) is the choice of
 * xg_amvp_select among them with quarter-pel AMVR (imvIdx 0): the prediction
 * SAD of each candidate against the block in the frame's reference.  The MV
 * and predictor fields of submitted blocks are ignored. */
typedef struct xg_server xg_server;

typedef struct xg_server_cfg {
    uint64_t      hbmBudget;       /* bytes of stream state the server may hold      */
    int32_t       maxStreams;      /* <= 32                                            */
    int32_t       maxBatchBlocks;  /* blocks per batch.  A batch waits for the other
                                      session's batch only if they share a stream,
                                      so a cap that splits the streams in two lets
                                      both sessions' batches run side by side      */
    int32_t       searchRange;     /* integer search range in pel (<= 32)              */
    uint32_t      searchLambda;    /* sadLambda of the integer search                  */
    int32_t       maxDelayUs;      /* a free session waits up to this long after the
                                      oldest waiting frame arrived, to gather more
                                      frames into its batch (0 = launch at once)      */
    xg_refine_cfg refine;          /* fractional refinement + MC                       */
    int32_t       minBatchFrames;  /* a free session holds its batch until this many
                                      streams (at most the open ones) have a frame to
                                      launch; bounded by maxDelayUs when that is set
                                      (0 = no minimum)                                 */
} xg_server_cfg;

typedef struct xg_frame_result {
    int32_t  stream;               /* stream id                                        */
    int32_t  status;               /* XG_OK; XG_ERR_STATE if dropped by close;
                                      XG_ERR_HIP if its batch failed on the device    */
    int32_t  nblocks;
    int32_t  batchFrames;          /* frames in the batch this frame ran in            */
    uint64_t tag;                  /* caller's tag from the submit call                */
    int64_t  submitNs, startNs, doneNs;   /* steady-clock times                       */
} xg_frame_result;

typedef struct xg_server_stats {
    int64_t  framesDone, framesDropped, batches;
    uint64_t bytesInUse, bytesBudget;
    int32_t  streamsOpen;
} xg_server_stats;

/* Where the memory behind the pointers of a submitted frame lives. */
typedef enum xg_mem_kind {
    XG_MEM_HOST = 0,         /* pageable host memory: planes are copied before the
                                submit call returns; outputs are written when the
                                frame completes                                       */
    XG_MEM_HOST_PINNED = 1,  /* page-locked host memory (hipHostMalloc): read and
                                written by DMA; must stay valid and unchanged until
                                the frame is reported                                  */
    XG_MEM_DEVICE = 2        /* device memory on the server's GPU: same lifetime rule */
} xg_mem_kind;

typedef struct xg_server_frame {
    int32_t          planesKind;   /* xg_mem_kind of planes[]                          */
    const xg_pixel  *planes[3];    /* Y, Cb, Cr: unpadded 10-bit samples               */
    intptr_t         strides[3];   /* in samples                                       */
    const xg_block  *blocks;       /* host partition (x, y, w, h used), or NULL to reuse
                                      the stream's last one                           */
    int32_t          nblocks;
    int32_t          outputsKind;  /* xg_mem_kind of results and pred[]                */
    xg_block_result *results;      /* nblocks entries                                  */
    xg_pixel        *pred[3];      /* prediction planes, or pred[0] NULL for none      */
    intptr_t         predStride[3];/* in samples (0: the plane width)                  */
    uint64_t         tag;
} xg_server_frame;

int32_t xg_server_create(xg_server **out, const xg_server_cfg *cfg);
int32_t xg_server_destroy(xg_server *srv);
/* Admission by HBM budget: XG_ERR_BUDGET if the stream's state (two padded
 * pictures, one prediction picture, MV field, and block, task and output
 * buffers sized for a partition of 4x4 blocks) does not fit. */
int32_t xg_server_open_stream(xg_server *srv, int32_t width, int32_t height, int32_t *streamId);
/* Waits for the stream's frame in flight, drops its pending frames (reported
 * with status XG_ERR_STATE) and releases its memory. */
int32_t xg_server_close_stream(xg_server *srv, int32_t streamId);
/* Queues one frame.  Nothing is copied when planes are pinned or on the
 * device, and a partition equal to the stream's last is neither uploaded nor
 * re-assigned.  A new partition is validated (sides 4..64, powers of two, on
 * the 4-sample grid, inside the picture) and assigned to the persistent
 * workgroups on the calling thread. */
int32_t xg_server_submit_frame(xg_server *srv, int32_t streamId, const xg_server_frame *frame);
/* xg_server_submit_frame with pageable planes (copied before returning) and
 * pageable outputs: per-block results to `results` (nblocks entries,
 * caller-owned until the frame is reported) and, when `pred` is not NULL, the
 * prediction planes to pred[0..2] (strides predStride[]). */
int32_t xg_server_submit(xg_server *srv, int32_t streamId, const xg_pixel *const planes[3],
                         const intptr_t strides[3], const xg_block *blocks, int32_t nblocks,
                         xg_block_result *results, xg_pixel *const pred[3], const intptr_t predStride[3],
                         uint64_t tag);
/* Up to `cap` completed frames; waits up to timeoutMs for the first one.
 * Returns the number written (>= 0) or a negative xg_status. */
int32_t xg_server_poll(xg_server *srv, xg_frame_result *out, int32_t cap, int32_t timeoutMs);
int32_t xg_server_stats_get(xg_server *srv, xg_server_stats *stats);

/* Integer full search for a grid of bw x bh blocks over a whole frame
 * (device planes; ref padded by >= range + 4).  Cost and order as
 * Xg26xMotionEstFull with predictor (0,0) and the CU MV window of
 * h266_enc_init.c:1462-1465.  Outputs per block (raster order): best MV in
 * quarter pel (int16 x,y pairs), cost and SAD.  Block widths 8, 16 and 32 on a
 * dword-aligned current plane take the packed-SAD kernel; other widths (4..64)
 * a scalar one with the same results. */
int32_t xg_full_search_frame(xg_session *s, const xg_plane *cur, const xg_plane *ref,
                             int32_t bw, int32_t bh, int32_t range, uint32_t sadLambda,
                             int32_t ctuSize, int16_t *bestMv, uint32_t *bestCost,
                             uint32_t *bestSad);

/* Lookahead motion estimation on half-resolution pictures: the role of the
 * closed Xin26xLookaheadFrame (h26x_core/include/h26x_look_ahead.h:27), whose
 * per-unit MVs laMv start the inter CU search as laMv << 1
 * (h266_encoder/source/h266_analyze_inter_cu.c:247-261).  Both luma planes
 * (device; ref padded by an even pad >= 2 * (range + 4)) are downscaled 2x
 * into session buffers, each sample the rounded mean (a + b + c + d + 2) >> 2
 * of its 2x2 block (the reference over its padding too); then every
 * unitSize x unitSize unit (8 or 16) wholly inside the half-resolution
 * picture gets the integer full search +/-range of xg_full_search_frame,
 * predictor (0, 0), SAD + lambda * MV bits.  Outputs per unit, raster order:
 * laMv in half-resolution quarter pel (int16 x, y), cost, SAD (device). */
int32_t xg_lookahead_me(xg_session *s, const xg_plane *cur, const xg_plane *ref, int32_t unitSize,
                        int32_t range, uint32_t sadLambda, int16_t *laMv, uint32_t *cost, uint32_t *sad);

/* xg_lookahead_me without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, checks and outputs as xg_lookahead_me, but no session: the
 * kernels are enqueued on `stream` (a hipStream_t; NULL is the null stream)
 * and both half-resolution pictures live in a static device array of
 * XG_LOOKAHEAD_STATIC_MAX_BYTES that every device the library runs on holds
 * until the process ends.  The two pictures need
 * 2 * (align128(even(w/2) * (h/2)) + (w/2 + pad) * (h/2 + pad)) bytes for a
 * w x h current plane and a reference padded by `pad`; a 3840x2160 pair with
 * the +/-64 minimum padding of 136 needs 9.1 MB.  Larger inputs return
 * XG_ERR_ARG before anything is enqueued.  The call allocates nothing and
 * makes no synchronous HIP call; the scratch is fully rewritten before it is
 * read, so no state carries over between calls.  Calls whose work overlaps on
 * one device share that device's array and race: serialise them (one stream,
 * or events). */
#define XG_LOOKAHEAD_STATIC_MAX_BYTES (12u << 20)
int32_t xg_lookahead_me_static(void *stream, const xg_plane *cur, const xg_plane *ref, int32_t unitSize,
                               int32_t range, uint32_t sadLambda, int16_t *laMv, uint32_t *cost, uint32_t *sad);

/* Motion estimation of the motion-compensated temporal filter: the role of
 * the closed Xin26xMctfMotionEstFrame (h26x_core/include/h26x_mctf.h:27, per
 * reference, search level and block size) and Xin26xMctfSubSample (:35),
 * filling MV fields as l0Mv / l1Mv / l2Mv / mv of h26x_mctf_struct.h:117-127.
 * A design of this library, hierarchical as the published temporal filter of
 * VVC's reference software: 2x2-mean subsampling; per level and block, the
 * 3x3 lower-level MVs around (bx / 2, by / 2) scaled by the resolution factor
 * and the zero MV, then the integer positions +/-range around the best, then
 * at the finest level 7x7 grids of quarter-pel and 1/16-pel steps around the
 * best, the first strictly lowest SSD winning each comparison (rows before
 * columns); fractional predictions by the standard's 8-tap luma filter.  A
 * position whose 8-tap reach [pos - 3, pos + blockSize + 4] leaves the
 * reference's padding is not costed.  Planes are device planes; all work is
 * enqueued on the session's stream. */

/* dst (src / 2 in both sides, dst->pad samples of edge replication filled
 * too) = rounded 2x2 means (a + b + c + d + 2) >> 2 of src. */
int32_t xg_mctf_subsample(xg_session *s, const xg_plane *src, const xg_plane *dst);

/* One level over every blockSize x blockSize block (8 or 16) wholly inside
 * cur: range 0..8; lower (device int32 (x, y) pairs, lowerBw x lowerBh blocks,
 * 1/16 of the lower level's pel) may be NULL (no candidates: the search is
 * centred on zero); factor scales the lower MVs to this level; frac adds the
 * fractional rounds.  Outputs (device, raster order of the block grid): mv
 * as int32 (x, y) pairs in 1/16 of this level's pel, err the SSD. */
int32_t xg_mctf_me_level(xg_session *s, const xg_plane *cur, const xg_plane *ref, int32_t blockSize, int32_t range,
                         const int32_t *lower, int32_t lowerBw, int32_t lowerBh, int32_t factor, int32_t frac,
                         int32_t *mv, uint32_t *err);

/* The whole search of one frame pair (sides >= 64): the 4x level (16x16
 * blocks, +/-8), the 2x level (16x16, +/-5, lower MVs x 2), full resolution
 * (16x16, +/-5, x 2) and 8x8 blocks at full resolution (range 0, x 1,
 * fractional rounds).  The subsampled pictures (padded by 32 and 16) and the
 * coarser fields live in session buffers.  With the full-resolution reference
 * padded by at least 56 no position the hierarchy can reach is skipped.
 * Outputs: the 8x8 grid's mv (int32 pairs, 1/16 pel) and err (device). */
int32_t xg_mctf_me(xg_session *s, const xg_plane *cur, const xg_plane *ref, int32_t *mv, uint32_t *err);

/* AMVP predictor choice, batched: the static Xin266SelectMvp
 * (h266_encoder/source/h266_analyze_inter_cu.c:60-131, called at :1003) for
 * many blocks at once, with the rounding Xin266FillAmvpCand ends with
 * (h266_get_neighbour_mv.c:2549-2558, Xin266ChangeMv32Prec :212-243) and the
 * ME-precision predictor of Xin266ChangeMv2Me (:293-320).  Per block, each
 * candidate (1/16 pel) is rounded to the AMVR precision of imvIdx (0 quarter,
 * 1 integer, 2 4-pel, 3 half: amvrPrecision, h266_entropy_write.c:120-123)
 * and back, its luma prediction formed as Xin266LumaInterPred does
 * (h266_inter_prediction.c:302-354: MV clipped to the CU window of
 * h266_enc_init.c:1462-1465, 8-tap filterIndex 0, the alternative half-pel
 * filter when imvIdx is 3) and compared with the block by SAD
 * (pfXinComputeDist); the first strictly lowest wins. */
typedef struct xg_amvp_block {
    int16_t x, y;          /* luma position                                  */
    uint8_t w, h;          /* 4..64, multiples of 4                          */
    uint8_t pic, ref;      /* session picture slots: current / reference     */
    int32_t ncand;         /* 1 or 2 candidates                              */
    int32_t cand[2][2];    /* candidate MVs, 1/16 pel, before AMVR rounding  */
} xg_amvp_block;

typedef struct xg_amvp_result {
    int32_t  mvpIdx;       /* chosen candidate; -1: the block was rejected   */
    uint32_t sad;          /* its prediction SAD (mvpSad)                    */
    int32_t  mv[2];        /* the chosen candidate after rounding, 1/16 pel  */
    int16_t  meMv[2];      /* the same in ME precision (quarter pel)         */
} xg_amvp_result;

/* Blocks and results are device arrays; the work is enqueued on the session's
 * stream (xg_session_sync waits for it).  A block outside its picture, with a
 * side outside 4..64 or not a multiple of 4, an unset slot or a reference
 * padded by less than ctuSize + 12 gets mvpIdx -1 and sad 0xFFFFFFFF; the
 * others are unaffected. */
int32_t xg_amvp_select(xg_session *s, const xg_amvp_block *blocks, int32_t n, int32_t imvIdx, int32_t ctuSize,
                       xg_amvp_result *out);

/* Candidate lists of xg_amvp_select from an MV field of 4x4 units (device,
 * int32 (x, y) pairs in 1/16 pel, fieldStride units per row, covering a
 * width x height picture): the rule of Xin266FillAmvpCand
 * (h266_get_neighbour_mv.c:2451-2560) with every unit inside the picture an
 * available inter neighbour of the same reference and no temporal or history
 * candidates: the below-left (A0) else left (A1) unit, the above-right (B0)
 * else above (B1) else above-left (B2) unit, a duplicate pair reduced to one,
 * a zero MV added when the list is empty or its one entry is non-zero, two
 * candidates only when they differ.  blocks (device) give positions, sizes
 * and slots; out (device) receives the lists. */
int32_t xg_amvp_from_field(xg_session *s, const int32_t *field, int32_t fieldStride, int32_t width, int32_t height,
                           const xg_block *blocks, int32_t n, xg_amvp_block *out);

/* Full motion-search pipeline of one frame with two reference pictures, on a
 * raster grid of 16x16 blocks (frame sides multiples of 16): per block and
 * reference, integer full search +/-range around (0,0) (as Xg26xMotionEstFull,
 * predictor (0,0)); predictor = component-wise median of the left, top and
 * top-right blocks' full-search MVs; start candidates as Xin266FillMeCand
 * (h266_analyze_inter_cu.c:133-293) from that predictor and the own, left,
 * top and top-right MVs; diamond search +/-range around the rounded predictor
 * (as Xg26xMotionEstDia); half- then quarter-pel rings (Xin266FullSubPelMe,
 * h266_sub_pel_refine.c:81-212, filterIndex 0); then the reference with the
 * lower quarter-pel cost (ties: ref0) is refined at 1/16 pel by
 * xg_refine_frac's kernel (candidate offsets around 4x the quarter-pel MV,
 * predictor 4x the median) and its prediction written to `pred`.
 * Pictures are installed in session slots 0 (cur), 1 (ref0) and 2 (ref1); the
 * session's block list becomes the grid.  mode: XG_MODE_BATCHED (kernels
 * enqueued in order) or XG_MODE_GRAPH (the same captured once in one HIP
 * graph and replayed).  Results: xg_session_read / xg_session_outputs for the
 * refinement, xg_me_pipeline_read for the earlier stages. */
typedef struct xg_pipeline_cfg {
    int32_t       range;           /* integer search range in pel (<= 64)          */
    uint32_t      sadLambda;       /* integer and quarter-pel stages               */
    int32_t       ctuSize;         /* CU MV window (h266_enc_init.c:1462-1465)     */
    xg_refine_cfg refine;          /* 1/16-pel stage and motion compensation        */
} xg_pipeline_cfg;

int32_t xg_me_pipeline(xg_session *s, const xg_pipeline_cfg *cfg, const xg_picture *cur,
                       const xg_picture *ref0, const xg_picture *ref1, const xg_picture *pred,
                       int32_t mode);
/* Host copies of the stage outputs of the last pipeline run, per block b and
 * reference r (any pointer may be NULL): fullMv / intMv [(b*2+r)*2 + {x,y}] in
 * integer pel, predMv / qMv likewise in quarter pel, intCost / qCost [b*2+r]. */
int32_t xg_me_pipeline_read(xg_session *s, int16_t *fullMv, int16_t *predMv, int16_t *intMv,
                            uint32_t *intCost, int16_t *qMv, uint32_t *qCost);

/* Batched forward transform of residual blocks on int8 MFMA.
 * xg_fdct_plan takes HOST arrays: block i's residual is at resid + offset[i]
 * (int16, row-major, stride = width; offset a multiple of min(8, width)
 * samples, so rows move as whole vectors) and its coefficients go to
 * coeff + offset[i] (int32, row-major [ver freq][hor freq], stride = width);
 * desc[i] = lgW | lgH << 4 | mtsIdx << 8 (sides 2..64; mtsIdx 0 or 2..5 as
 * h26x_trans_context.h:42-51, DST-VII/DCT-VIII only for sides 4..32).  The plan
 * groups blocks by shape and stays on the device until replaced.
 * xg_fdct_run enqueues the transform of every planned block on the session
 * stream (device pointers).  Shifts for 10-bit samples as Xg266FDctWxH;
 * coefficients past the VVC zero-out are written as 0. */
int32_t xg_fdct_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, int32_t nblocks);
int32_t xg_fdct_run(xg_session *s, const int16_t *resid, xg_coeff *coeff);

/* Batched reconstruction: inverse transform of dequantised coefficients and
 * the prediction add, for TUs whose inverse does not depend on each other
 * (the batched form of Xin266IDctAdd, h266_encoder/source/h266_idct_add.c:45-104,
 * as Xin266ReconCu calls it per TU at h266_itrans_recon.c:127-199; the LMCS
 * chroma scaling is not applied).
 * xg_recon_plan takes HOST arrays: block i's coefficients are at
 * coeff + offset[i] (int16 -- the dequantiser's output range --, row-major
 * [ver freq][hor freq], stride = width), its prediction at pred + offset[i]
 * and its reconstruction goes to recon + offset[i] (10-bit samples, stride =
 * width); offsets are multiples of min(8, width) samples.
 * desc[i] = lgW | lgH << 4 | mtsIdx << 8 | implicitMts << 12: sides 2..64,
 * mtsIdx 0..5 as h26x_trans_context.h:42-51 (1 = transform skip), and
 * implicitMts = 1 for an intra luma TU with MTS enabled, which turns DCT-II
 * sides of 4..16 into DST-VII when mtsIdx is 0 (h266_inverse_trans.c:149-159).
 * Coefficients past the zero-out (32 for DCT-II-64, 16 for 32-point
 * DST-VII/DCT-VIII) are not read.  Shifts 7 and 10 (10-bit samples).
 * xg_recon_run enqueues every planned block on the session stream (device
 * pointers); recon must not overlap coeff or pred. */
int32_t xg_recon_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, int32_t nblocks);
int32_t xg_recon_run(xg_session *s, const int16_t *coeff, const xg_pixel *pred, xg_pixel *recon);

/* Batched quantisation + dequantisation: pfXinQuantInvQuant for every TU of a
 * batch, as Xin266Transform calls it (h266_encoder/source/h266_trans_quant.c:
 * 204-222) with the parameters it derives at :104-120 from the tables of
 * h266_encoder/source/h266_encoder_create.c:94-104,1677-1705 (10-bit); the
 * per-TU arithmetic of Xin266QuantInvQuant (h266_quant_inv_quant.c:80-143).
 * xg_quant_plan takes HOST arrays: TU i's forward coefficients are at
 * tCoeff + offset[i] (xg_coeff, row-major, stride = width; offset a multiple of
 * min(4, width)) and its levels and dequantised coefficients go to the same
 * offset of qCoeff / rCoeff (int16 -- both are clipped to that range --, the
 * layout xg_recon_run reads).  desc[i] = lgW | lgH << 4 | mtsIdx << 8 |
 * lgCGW << 12 | lgCGH << 15 | qp << 18 | rdoq << 24 | intraFrame << 25:
 * sides 2..64; mtsIdx 1 = transform skip; coefficient groups of
 * 2^lgCGW x 2^lgCGH inside the min(W, 32) x min(H, 32) region that is
 * quantised (at most 64 groups there, 256 in the TU, 64 coefficients each);
 * qp the quantParam index (QP + 12 for 10-bit, 0..63); rdoq selects the
 * 1 << (qShift - 1) rounding of an RDOQ configuration, else the dead zone of
 * 171 (intraFrame, frameType >= I) or 85 << (qShift - 9).  Past the region
 * both outputs are zero.
 * xg_quant_run enqueues every planned TU on the session stream (device
 * pointers: tCoeff 16-byte aligned, the others 8-byte aligned) and writes
 * nzMap[i]: bit k set when group k of TU i's region (raster order) holds a
 * non-zero level, as *nzCGBitMapRs. */
int32_t xg_quant_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, int32_t nblocks);
int32_t xg_quant_run(xg_session *s, const xg_coeff *tCoeff, int16_t *qCoeff, int16_t *rCoeff, uint64_t *nzMap);

/* Forward transform with the quantiser fused into its last stage: the
 * Xin266SubFDctWxH -> pfXinQuantInvQuant sequence of Xin266Transform
 * (h266_encoder/source/h266_trans_quant.c:146-222) for a batch of TUs, with
 * no int32 coefficient buffer in between.  xg_fdct_quant_plan takes the
 * arguments of xg_fdct_plan plus qdesc[i], block i's xg_quant_plan descriptor:
 * its sides and mtsIdx must repeat desc[i]'s and its coefficient groups must
 * be 4x4 (sides >= 4).  xg_fdct_quant_run writes exactly what xg_fdct_run
 * followed by xg_quant_run writes to qCoeff, rCoeff and nzMap (device
 * pointers, 8-byte aligned); it counts as the forward-transform stage (2) of
 * the session timers.  A plan made by xg_fdct_plan serves xg_fdct_run only;
 * one made by xg_fdct_quant_plan serves both. */
int32_t xg_fdct_quant_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, const uint32_t *qdesc,
                           int32_t nblocks);
int32_t xg_fdct_quant_run(xg_session *s, const int16_t *resid, int16_t *qCoeff, int16_t *rCoeff, uint64_t *nzMap);

/* Batched DMVR (the refinement of Xg266DeriveMvFromDmvr for many blocks).
 * xg_dmvr_plan takes a HOST array of blocks and writes subOffset[0..n]
 * (optional): block i's subblock deltas are mvd[2*subOffset[i] ..
 * 2*subOffset[i+1]).  xg_dmvr_run enqueues the refinement of every planned
 * block on the session stream: ref0 / ref1 are device luma planes (data at
 * sample (0, 0), padded by `pad` samples), and every block's reads (its
 * clipped MVs, -2 pels, w + 5 x h + 5 samples) must lie inside the padding,
 * else XG_ERR_ARG.  mvd is a device array of int16 (x, y) pairs. */
typedef struct xg_dmvr_block {
    int32_t x, y, width, height;                 /* luma position, sides 8..128 */
    int32_t mv0x, mv0y, mv1x, mv1y;              /* 1/16 pel */
    int16_t minMvX, minMvY, maxMvX, maxMvY;      /* integer-pel MV window */
} xg_dmvr_block;
int32_t xg_dmvr_plan(xg_session *s, const xg_dmvr_block *blocks, int32_t nblocks, int32_t *subOffset);
int32_t xg_dmvr_run(xg_session *s, const xg_plane *ref0, const xg_plane *ref1, int16_t *mvd);

/* Bi-prediction distortion of a frame's blocks: the B-frame step of
 * h266_encoder/source/h266_analyze_inter_cu.c:1212-1222, pfXinComputeAvgDist
 * (input, pred0, pred1) for every block in one launch -- SATD (satd != 0, as
 * Xg266ComputeAvgSatd: 8x8 Hadamard tiles when both sides are multiples of 8,
 * else 4x4) or SAD of input - ((pred0 + pred1 + 1) >> 1).
 * xg_bipred_plan takes a HOST array of blocks (x, y, w, h; sides multiples of
 * 4 in 4..128); xg_bipred_run checks every block against the three device
 * planes and writes dist[i] (device) for block i. */
int32_t xg_bipred_plan(xg_session *s, const xg_block *blocks, int32_t nblocks);
int32_t xg_bipred_run(xg_session *s, const xg_plane *input, const xg_plane *pred0, const xg_plane *pred1,
                      int32_t satd, uint32_t *dist);

#ifdef __cplusplus
}
#endif
#endif

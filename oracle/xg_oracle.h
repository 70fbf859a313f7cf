/*
 * xg_oracle.h -- CPU oracle for the VVC block-matching hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product library (libxg266.so) never links, loads or calls it.
 *
 * Every function restates the reference algorithm in plain scalar C.  Each
 * definition in xg_oracle.c cites the reference file:line it follows (paths
 * relative to the xin26x source tree).  Where the reference implementation
 * is a closed binary, the function states the public method it follows and
 * is marked "parity unpinned vs the reference"; such functions are pinned
 * by closed forms, invariants and brute force in tests/.
 *
 * 10-bit build: PIXEL = uint16 in [0,1023], COEFF = int32.
 */
#ifndef XG_ORACLE_H
#define XG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint16_t or_pixel;
typedef int32_t  or_coeff;

/* ---- T1: luma 8-tap interpolation (h266_core/source/h266_luma_mc.c) ---- */
void or_InterpCopy(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                   int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_LumaInterpHor(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                      int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_LumaInterpVet(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                      int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_LumaInterpHorVet(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                         int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_InterpCopyU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                        int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_LumaInterpHorU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                           int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_LumaInterpVetU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                           int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_LumaInterpHorVetU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                              int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_LumaInterpVetS16U8(const int16_t *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                           int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_LumaInterpVetS16S16(const int16_t *src, intptr_t ss, int16_t *dst, intptr_t ds,
                            int32_t frac, uint32_t w, uint32_t h, uint32_t fi);

/* ---- chroma 4-tap interpolation (h266_core/source/h266_chroma_mc.c) ---- */
void or_ChromaInterpHor(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                        int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_ChromaInterpVet(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                        int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_ChromaInterpHorVet(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                           int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_ChromaInterpHorU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                             int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_ChromaInterpVetU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                             int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_ChromaInterpHorVetU8S16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                                int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_ChromaInterpVetS16U8(const int16_t *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                             int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
void or_ChromaInterpVetS16S16(const int16_t *src, intptr_t ss, int16_t *dst, intptr_t ds,
                              int32_t frac, uint32_t w, uint32_t h, uint32_t fi);

/* ---- bi-prediction combine (h266_luma_mc.c:561-629) ---- */
void or_InterpAvgS16U8(const int16_t *s0, intptr_t st0, const int16_t *s1, intptr_t st1,
                       or_pixel *dst, intptr_t ds, uint32_t w, uint32_t h);
void or_InterpWeightS16U8(const int16_t *s0, intptr_t st0, const int16_t *s1, intptr_t st1,
                          or_pixel *dst, intptr_t ds, uint32_t w, uint32_t h,
                          int32_t wA, int32_t wB);

/* ---- table dispatch as the encoder does it ---- */
/* h266_inter_prediction.c:143-172 (luma, PIXEL out, table index [!!fx][!!fy]) */
void or_LumaInterpolation(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                          int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
/* h266_inter_prediction.c:78-141 (luma, S16 out) */
void or_LumaInterpolationS16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                             int32_t frac, uint32_t w, uint32_t h, uint32_t fi);
/* h266_inter_prediction.c:174-236 (chroma, PIXEL out; frac = (fy<<5)|fx) */
void or_ChromaInterpolation(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds,
                            int32_t frac, uint32_t w, uint32_t h);
/* h266_inter_prediction.c:238-300 (chroma, S16 out) */
void or_ChromaInterpolationS16(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds,
                               int32_t frac, uint32_t w, uint32_t h);

/* ---- distortion (closed in the reference: closed forms / HM methods) ---- */
uint32_t or_ComputeSad(const or_pixel *in, intptr_t is, const or_pixel *pr, intptr_t ps,
                       uint32_t w, uint32_t h);
void or_ComputeSadx8(const or_pixel *in, intptr_t is, const or_pixel *const ref[8], intptr_t rs,
                     uint32_t w, uint32_t h, uint32_t sad[8]);
void or_ComputeSadx5(const or_pixel *in, intptr_t is, const or_pixel *r0, const or_pixel *r1,
                     const or_pixel *r2, const or_pixel *r3, const or_pixel *r4, intptr_t rs,
                     uint32_t w, uint32_t h, uint32_t sad[5]);
void or_ComputeSadx3(const or_pixel *in, intptr_t is, const or_pixel *r0, const or_pixel *r1,
                     const or_pixel *r2, intptr_t rs, uint32_t w, uint32_t h, uint32_t sad[3]);
uint32_t or_ComputeSatd(const or_pixel *in, intptr_t is, const or_pixel *pr, intptr_t ps,
                        uint32_t w, uint32_t h);
uint32_t or_ComputeHada8x8(const or_pixel *in, intptr_t is);
/* bi-prediction distortions against (a + b + 1) >> 1 (h266_func_struct.h:65-73, 96-104) */
uint32_t or_ComputeAvgSad(const or_pixel *in, intptr_t is, const or_pixel *a, const or_pixel *b, intptr_t rs,
                          uint32_t w, uint32_t h);
uint32_t or_ComputeAvgSatd(const or_pixel *in, intptr_t is, const or_pixel *a, const or_pixel *b, intptr_t rs,
                           uint32_t w, uint32_t h);
void or_ComputeAvgSadx8(const or_pixel *in, intptr_t is, const or_pixel *const a[8], const or_pixel *const b[8],
                        intptr_t rs, uint32_t w, uint32_t h, uint32_t sad[8]);

/* ---- DMVR (h266_bilinear_filter.c; h266_inter_prediction.c:864-1195) ---- */
void or_BiliInterpCopy(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                       uint32_t w, uint32_t h);
void or_BiliInterpHor(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                      uint32_t w, uint32_t h);
void or_BiliInterpVet(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                      uint32_t w, uint32_t h);
void or_BiliInterpHorVet(const or_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                         uint32_t w, uint32_t h);
uint32_t or_ComputeSadS16(const int16_t *a, intptr_t as, const int16_t *b, intptr_t bs, uint32_t w, uint32_t h,
                          uint32_t subShift);
typedef struct or_dmvr {
    const or_pixel *ref0; intptr_t stride0;     /* sample (0, 0) of each padded reference plane */
    const or_pixel *ref1; intptr_t stride1;
    int32_t  cuPelX, cuPelY;                    /* block position, luma samples */
    uint32_t width, height;
    int32_t  mv0x, mv0y, mv1x, mv1y;            /* 1/16 pel */
    int16_t  minMvX, minMvY, maxMvX, maxMvY;    /* integer-pel MV window */
} or_dmvr;
void or_DeriveMvFromDmvr(const or_dmvr *d, int16_t *mvd);
/* the parametric sub-pel step: costs = centre, left, top, right, bottom; dmv in 1/16 pel */
void or_DmvrSubPel(int non0Cost, int16_t *dmv, const int64_t costs[5]);

/* ---- lookahead: 2x2 rounded-mean downscale (a design choice; the reference's is closed) ---- */
void or_Downscale2x2(const or_pixel *src, intptr_t ss, or_pixel *dst, intptr_t ds, uint32_t w, uint32_t h);

/* ---- rate terms ---- */
uint32_t or_ExpGolombBits(int32_t v);
uint32_t or_MvdBits(int32_t dx, int32_t dy);
uint32_t or_SadLambda(int32_t qp, int32_t is10bit);

/* ---- MV helpers (h266_get_neighbour_mv.c:245-320) ---- */
int32_t or_MvRoundShift(int32_t v, int32_t shift);

/* ---- integer and fractional motion search ---- */
typedef struct or_me {
    uint32_t width, height;
    int16_t  minMvX, minMvY, maxMvX, maxMvY;   /* int-pel */
    int16_t  candX[12], candY[12];             /* int-pel start candidates */
    uint32_t candNum;
    int16_t  bestMvX, bestMvY;                 /* quarter-pel */
    int16_t  predMvX, predMvY;                 /* quarter-pel */
    uint32_t bestCost, bestSad;
    int32_t  searchWidth, searchHeight;
    uint32_t sadLambda;
    const or_pixel *input; intptr_t inputStride;
    const or_pixel *ref;   intptr_t refStride; /* co-located block in padded ref */
} or_me;

void or_MotionEstFull(or_me *me);
void or_MotionEstDia(or_me *me);
void or_FullSubPelMe(or_me *me);
void or_FillMeCand(int16_t predX, int16_t predY, const int16_t *qx, const int16_t *qy,
                   uint32_t n, int16_t *outX, int16_t *outY, uint32_t *outNum);

/* ---- forward transforms (closed in the reference; H.266 matrices) ---- */
enum { OR_DCT2 = 0, OR_DCT8 = 1, OR_DST7 = 2, OR_SKIP = 4 };
int  or_TrMatrix(int type, int n, int32_t *out);
void or_FDct1d(int type, int n, const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
               int32_t line, int32_t shift);
void or_FDct2WxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
                 uint32_t w, uint32_t h, or_coeff *tmp, int32_t s1, int32_t s2);
void or_BlockSubForDct(const or_pixel *in, intptr_t is, const or_pixel *pred, intptr_t ps,
                       or_coeff *diff, intptr_t dstride, uint32_t w, uint32_t h);
void or_FDctWxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
                uint32_t w, uint32_t h, uint32_t mtsIdx, int isIntraLuma, int enableMts,
                int bitDepth);
void or_MtsTypes(uint32_t mtsIdx, int32_t *horTx, int32_t *verTx);

/* ---- inverse transforms and reconstruction ---- */
void or_IDct1d(int type, int n, const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
               int32_t line, int32_t skipLine, int32_t shift);
void or_IDct2WxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
                 uint32_t w, uint32_t h, or_coeff *tmp, int32_t s1, int32_t s2);
void or_ISkipWxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os, uint32_t w, uint32_t h);
int  or_IDctWxH(const or_coeff *in, intptr_t is, or_coeff *out, intptr_t os,
                uint32_t w, uint32_t h, uint32_t mtsIdx, int isIntraLuma, int enableMts, int bitDepth);
void or_BlockRecon(const or_coeff *res, intptr_t rs, const or_pixel *pred, intptr_t ps,
                   or_pixel *recon, intptr_t os, uint32_t w, uint32_t h, int bitDepth);
void or_ReconBatch(const int16_t *coeff, const or_pixel *pred, const int64_t *offset, const uint32_t *desc,
                   int32_t i0, int32_t i1, or_pixel *recon);
int  or_IDctAdd(const or_coeff *in, intptr_t is, const or_pixel *pred, intptr_t ps, or_pixel *out,
                intptr_t os, uint32_t w, uint32_t h, uint32_t mtsIdx, int isIntraLuma, int enableMts,
                int bitDepth);
/* Blocks [i0, i1) of a batch in the layout of the batched GPU entry
 * (int16 residuals at resid + offset[i], desc = lgW | lgH << 4 | mtsIdx << 8). */
void or_FDctBatch(const int16_t *resid, const int64_t *offset, const uint32_t *desc, int32_t i0, int32_t i1,
                  or_coeff *coeff);

/* ---- frame-level drivers used as the batched-path checker ---- */
typedef struct or_block {
    int32_t x, y, w, h;        /* luma position and size */
    int32_t mvx, mvy;          /* base MV, 1/16 pel */
    int32_t mvpx, mvpy;        /* predictor, 1/16 pel */
} or_block;

typedef struct or_refine_prm {
    int32_t ncand;
    const int16_t *offx, *offy;    /* candidate offsets, 1/16 pel */
    uint32_t lambda;
    int32_t  ctuSize;              /* MV clip window (h266_enc_init.c:1462-1465) */
    int32_t  useChroma;            /* add Cb/Cr SATD into the candidate cost */
} or_refine_prm;

/* Config-2 stage: per block and candidate, luma 8-tap (+ chroma 4-tap)
 * interpolation and SATD; returns per-candidate costs, the best candidate,
 * and the motion-compensated prediction of the best candidate. */
void or_RefineFracBlocks(const or_pixel *curY, const or_pixel *curU, const or_pixel *curV,
                         intptr_t curStrideY, intptr_t curStrideC,
                         const or_pixel *refY, const or_pixel *refU, const or_pixel *refV,
                         intptr_t refStrideY, intptr_t refStrideC,
                         int32_t frameW, int32_t frameH,
                         const or_block *blk, int32_t nblk, const or_refine_prm *prm,
                         uint32_t *candCost, int32_t *bestIdx, uint32_t *bestCost,
                         int32_t *bestMvx, int32_t *bestMvy,
                         or_pixel *predY, or_pixel *predU, or_pixel *predV,
                         intptr_t predStrideY, intptr_t predStrideC);

/* Full motion-search pipeline (two references, grid of bw x bh blocks):
 * full search, median predictor + FillMeCand + diamond, half/quarter-pel
 * ring, reference choice, 1/16-pel refinement + MC.  Layout in xg_oracle.c. */
/* Stage-2 predictor and start candidates of the pipeline for block b and
 * reference r (see or_MePipeline). */
void or_PipelineCands(const int16_t *fullMv, int32_t nbx, int32_t b, int32_t r, int16_t *predX, int16_t *predY,
                      int16_t *candX, int16_t *candY, uint32_t *candNum);
void or_MePipeline(const or_pixel *curY, const or_pixel *curU, const or_pixel *curV, intptr_t csy, intptr_t csc,
                   const or_pixel *r0Y, const or_pixel *r0U, const or_pixel *r0V,
                   const or_pixel *r1Y, const or_pixel *r1U, const or_pixel *r1V, intptr_t rsy, intptr_t rsc,
                   int32_t fw, int32_t fh, int32_t bw, int32_t bh, int32_t range, uint32_t sadLambda,
                   int32_t ctuSize, const or_refine_prm *prm,
                   int16_t *fullMv, int16_t *predMv, int16_t *intMv, uint32_t *intCost, int16_t *qMv,
                   uint32_t *qCost, int32_t *refIdx, uint32_t *candCost, int32_t *bestIdx, uint32_t *bestCost,
                   int32_t *bestMvx, int32_t *bestMvy, or_pixel *predY, or_pixel *predU, or_pixel *predV,
                   intptr_t psy, intptr_t psc);
/* or_MePipeline for the blocks sample[0..n) only, outputs indexed by sample
 * position (see xg_oracle.c); thread-safe across disjoint samples. */
void or_MePipelineSample(const or_pixel *curY, const or_pixel *curU, const or_pixel *curV, intptr_t csy,
                         intptr_t csc, const or_pixel *r0Y, const or_pixel *r0U, const or_pixel *r0V,
                         const or_pixel *r1Y, const or_pixel *r1U, const or_pixel *r1V, intptr_t rsy, intptr_t rsc,
                         int32_t fw, int32_t fh, int32_t bw, int32_t bh, int32_t range, uint32_t sadLambda,
                         int32_t ctuSize, const or_refine_prm *prm, const int32_t *sample, int32_t n,
                         int16_t *fullMv, int16_t *predMv, int16_t *intMv, uint32_t *intCost, int16_t *qMv,
                         uint32_t *qCost, int32_t *refIdx, uint32_t *candCost, int32_t *bestIdx, uint32_t *bestCost,
                         int32_t *bestMvx, int32_t *bestMvy, or_pixel *predY, or_pixel *predU, or_pixel *predV,
                         intptr_t psy, intptr_t psc);

/* Config-1 stage: integer full search for every block of a frame. */
/* Integer search of the serving path (predictor-centred, centre clipped into
 * the CU window) for blocks [i0, i1); writes each block's mvx/mvy. */
void or_SearchBlocks(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs, int32_t fw, int32_t fh,
                     or_block *blk, int32_t i0, int32_t i1, int32_t range, uint32_t sadLambda, int32_t ctuSize);
void or_FullSearchFrame(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs,
                        int32_t frameW, int32_t frameH, int32_t bw, int32_t bh,
                        int32_t range, uint32_t sadLambda, int32_t ctuSize,
                        int32_t blk0, int32_t blk1,
                        int16_t *bestMvX, int16_t *bestMvY, uint32_t *bestCost,
                        uint32_t *bestSad);

/* ---- AMVP predictor selection ------------------------------------------ */
/* One MV component from precision inPrec to outPrec (log2 of the units per
 * 4 pel: 0 4-pel, 2 integer, 3 half, 4 quarter, 6 sixteenth), as
 * Xin266ChangeMv32Prec (h266_get_neighbour_mv.c:212-243) does it. */
int32_t or_ChangeMvPrec(int32_t v, int32_t inPrec, int32_t outPrec);
/* AMVP list of the block (x, y, w, h) from an MV field of 4x4 units
 * (field[(uy * stride + ux) * 2 + {0,1}], 1/16 pel, every unit inside the
 * fw x fh picture an available inter neighbour): the rule of
 * Xin266FillAmvpCand (h266_get_neighbour_mv.c:2451-2560) without temporal
 * and history candidates.  Returns the number of candidates (1 or 2), written
 * to cand[0..1] / cand[2..3], before AMVR rounding. */
int32_t or_AmvpFromField(const int32_t *field, intptr_t stride, int32_t fw, int32_t fh, int32_t x, int32_t y,
                         int32_t w, int32_t h, int32_t cand[4]);
/* Xin266SelectMvp (h266_analyze_inter_cu.c:60-131) for one block: each of the
 * ncand candidates (1/16 pel) rounded to the AMVR precision of imvIdx
 * (0 quarter, 1 integer, 2 4-pel, 3 half; h266_entropy_write.c:120-123) and
 * back, its luma prediction (MV clipped to the CU window; the alternative
 * half-pel filter when imvIdx is 3) compared with the block by SAD; the
 * first strictly lowest wins.  Returns its index; sad, the rounded MV (1/16
 * pel) and the same in ME precision (quarter pel, Xin266ChangeMv2Me,
 * h266_get_neighbour_mv.c:293-320).  cur/ref point at sample (0,0). */
int32_t or_SelectMvp(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs, int32_t fw, int32_t fh,
                     int32_t x, int32_t y, int32_t w, int32_t h, int32_t ncand, const int32_t *cand,
                     int32_t imvIdx, int32_t ctuSize, uint32_t *sad, int32_t mv[2], int16_t meMv[2]);

/* ---- MCTF motion estimation (design of the library; see or_MctfMeLevel) -- */
/* dst (w/2 x h/2 plus `pad` samples of edge replication on every side) =
 * rounded 2x2 means (a + b + c + d + 2) >> 2 of src (w x h); dst points at
 * its sample (0, 0). */
void or_MctfSubsample(const or_pixel *src, intptr_t ss, int32_t w, int32_t h, or_pixel *dst, intptr_t ds,
                      int32_t pad);
/* One level of the MCTF motion estimation: for every bs x bs block wholly
 * inside the w x h picture (raster grid), the MV (1/16 of this level's pel)
 * and SSD of the best of: the MVs of the 3x3 lower-level blocks around
 * (bx / 2, by / 2) (lowerBw x lowerBh grid, times `factor`; lower NULL: none)
 * and zero, scanned rows then columns, then the integer positions +/-range
 * around that best (rows then columns), then, when frac is set, a 7x7 grid
 * of quarter-pel steps and a 7x7 grid of 1/16-pel steps around the best
 * (the standard's 8-tap luma prediction); the first strictly lowest SSD wins
 * each comparison.  A position whose 8-tap reach [pos - 3, pos + bs + 4]
 * leaves the reference's `pad` samples of padding is not evaluated.  Block
 * rows [by0, by1) are computed; mv and err hold nbx * nby entries (x, y pairs
 * for mv). */
void or_MctfMeLevel(const or_pixel *cur, intptr_t cs, const or_pixel *ref, intptr_t rs, int32_t w, int32_t h,
                    int32_t pad, int32_t bs, int32_t range, const int32_t *lower, int32_t lowerBw, int32_t lowerBh,
                    int32_t factor, int32_t frac, int32_t by0, int32_t by1, int32_t *mv, uint32_t *err);

/* ---- quantisation (h266_encoder/source/h266_quant_inv_quant.c:27-143) ---- */
void or_QuantInvQuant(int32_t *q, const int32_t *t, int32_t *r, intptr_t stride, uint32_t w, uint32_t h,
                      uint32_t cgw, uint32_t cgh, int32_t qMult, int32_t qAdd, int32_t qShift, int32_t iqMult,
                      int32_t iqAdd, int32_t iqShift, uint64_t *nzMap);
/* Per-TU quantiser parameters as Xin266Transform derives them
 * (h266_encoder/source/h266_trans_quant.c:104-120, tables of
 * h266_encoder/source/h266_encoder_create.c:94-104,1677-1705; 10-bit).
 * desc = lgW | lgH << 4 | mtsIdx << 8 | lgCGW << 12 | lgCGH << 15 | qp << 18
 * | rdoq << 24 | intraFrame << 25, qp the table index (QP + 12, 0..63),
 * mtsIdx 1 = transform skip.  out = {qMult, qAdd, qShift, iqMult, iqAdd,
 * iqShift}. */
void or_QuantParams(uint32_t desc, int32_t out[6]);
/* TUs [i0, i1) of a batch: TU i's coefficients at t + offset[i] (stride =
 * width); its levels and dequantised values go to q / r + offset[i] as int16,
 * over the min(w, 32) x min(h, 32) region (zeros past it), and its non-zero
 * group map to nzMap[i]. */
void or_QuantBatch(const int32_t *t, const int64_t *offset, const uint32_t *desc, int32_t i0, int32_t i1,
                   int16_t *q, int16_t *r, uint64_t *nzMap);

#ifdef __cplusplus
}
#endif
#endif

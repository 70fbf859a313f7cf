// xg_fdct.hip -- batched forward DCT-II / DST-VII / DCT-VIII on int8 MFMA.
//
// Per block (W x H residual, int16 in [-1023, 1023]) the two 1-D passes of
// h266_encoder/source/h266_forward_trans.c:112-129 are two integer matrix
// products computed with v_mfma_i32_16x16x32_i8:
//   stage 1  TMP[k][y] = (sum_n T_W[k][n] * r[y][n] + 2^(s1-1)) >> s1,  k < Wk
//   stage 2  OUT[j][k] = (sum_y T_H[j][y] * TMP[k][y] + 2^(s2-1)) >> s2, j < Hk
// with s1 = lgW + 1, s2 = lgH + 6 for 10-bit samples (:72-73) and the VVC
// zero-out (Wk, Hk = 32 for 64-point DCT-II, 16 for 32-point DST-VII/DCT-VIII).
// The basis matrices (entries in [-91, 91]) are int8 A operands.  The data are
// split into int8 limbs so every product is exact in int32:
//   r = 128*hi + lo, lo in [0,127], hi in [-8, 7]                (2 MFMAs)
//   t = 16384*l2 + 128*l1 + l0, l0,l1 in [0,127], |l2| <= 3      (3 MFMAs)
// |partial| <= 91*127*64 = 739,648; |stage-2 sum| <= 91*46.5K*64 < 2^31.
//
// Blocks of one descriptor are processed in groups of up to 4 (as many as
// fill a 16-wide MFMA tile: 16 / min(H, Wk)): stage 1 stacks the rows of all
// blocks of the group as the N dimension, stage 2 stacks their kept columns,
// so a 4x4 block costs a quarter of a tile instead of a whole one.
//
// Operand layout of v_mfma_i32_16x16x32_i8 (one wavefront): lane l supplies
// A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16] as 8 packed bytes and
// receives D[4*(l/16) + r][l%16], r = 0..3.
#include <algorithm>
#include <mutex>
#include <string.h>
#include <vector>

#include "xg_batch.hpp"
#include "xg_device.hpp"
#include "xg_quant.hpp"
#include "xg_runtime.hpp"
#include "xg_trmatrix.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));

namespace xg {

// Padded int8 matrices: one per (type, lgN); rows padded to max(16, N),
// columns to max(32, N), zeros in the padding.  For N <= 8 also a
// block-diagonal 16 x 32 form: 16 / N copies of the N x N matrix along the
// diagonal, so one MFMA tile transforms 16 / N row sets at once.
struct MatIndex { int off[3][7]; int cols[3][7]; int diag[3][4]; int total; };

static MatIndex mat_index()
{
    MatIndex m{};
    int o = 0;
    for (int t = 0; t < 3; t++)
        for (int lg = 1; lg <= 6; lg++) {
            const int N = 1 << lg;
            const bool ok = t == kDCT2 ? true : (N >= 4 && N <= 32);
            m.off[t][lg] = ok ? o : -1;
            m.cols[t][lg] = std::max(32, N);
            if (ok) o += std::max(16, N) * std::max(32, N);
        }
    for (int t = 0; t < 3; t++)
        for (int lg = 1; lg <= 3; lg++) {
            const bool ok = t == kDCT2 || lg >= 2;
            m.diag[t][lg] = ok ? o : -1;
            if (ok) o += 16 * 32;
        }
    m.total = o;
    return m;
}

static std::once_flag g_mat_once;
static int8_t *g_mats = nullptr;
static MatIndex g_mi;
static bool g_mat_ok = false;

static bool mats_init()
{
    std::call_once(g_mat_once, [] {
        g_mi = mat_index();
        std::vector<int8_t> h((size_t)g_mi.total, 0);
        for (int t = 0; t < 3; t++)
            for (int lg = 1; lg <= 6; lg++) {
                if (g_mi.off[t][lg] < 0) continue;
                const int N = 1 << lg, cols = g_mi.cols[t][lg];
                for (int k = 0; k < N; k++)
                    for (int n = 0; n < N; n++)
                        h[(size_t)g_mi.off[t][lg] + (size_t)k * cols + n] = (int8_t)tr_entry(kTr, t, N, k, n);
            }
        for (int t = 0; t < 3; t++)
            for (int lg = 1; lg <= 3; lg++) {
                if (g_mi.diag[t][lg] < 0) continue;
                const int N = 1 << lg;
                for (int q = 0; q < 16 / N; q++)
                    for (int k = 0; k < N; k++)
                        for (int n = 0; n < N; n++)
                            h[(size_t)g_mi.diag[t][lg] + (size_t)(q * N + k) * 32 + q * N + n] =
                                (int8_t)tr_entry(kTr, t, N, k, n);
            }
        if (!XG_HIP(hipMalloc((void **)&g_mats, h.size()))) return;
        if (!XG_HIP(hipMemcpy(g_mats, h.data(), h.size(), hipMemcpyHostToDevice))) return;
        g_mat_ok = true;
    });
    return g_mat_ok;
}

constexpr int kGroupMax = 64;     // blocks per group: one relative offset per lane

// LDS of a launch whose blocks are at most HMAX high: the group's residual
// (kStage samples) and its stage-1 output as three int8 limb planes of kPlane
// bytes each.  A limb plane holds one row of `tmp_stride(H)` bytes per kept
// column (b, k), the column's H values along the row.
template <int HMAX>
struct FdctLds {
    static constexpr int kStage = 64 * HMAX;
    static constexpr int kPlane = HMAX == 64 ? 32 * 72 : 2048;
};

__host__ __device__ inline int tmp_stride(int H)
{
    return H <= 8 ? 8 : H <= 16 ? 16 : H + 8;   // 8-byte aligned; long rows padded against bank conflicts
}

struct FdctGroup {
    long long base;     // element offset the group's relative block offsets start from
    uint32_t desc;      // lgW | lgH << 4 | horTx << 8 | verTx << 12
    int32_t count;      // blocks, 1..kGroupMax, all of this descriptor
    int32_t first;      // index of the group's first relative offset
    int32_t pad;
};

// Blocks of one descriptor per group: as many as the launch's LDS holds (the
// residual W*H int16 and Wk limb rows per block).
inline int group_cap(int W, int H, int Wk, int stage, int plane)
{
    int c = stage / (W * H);
    const int t = plane / (Wk * tmp_stride(H));
    if (t < c) c = t;
    return c < kGroupMax ? c : kGroupMax;
}

struct FdctArgs {
    const int16_t *resid;
    int32_t *coeff;
    int16_t *q, *r;                // fused quantiser: levels and dequantised values (int16)
    unsigned long long *nz;        // fused quantiser: group map per TU
    const uint32_t *qd;            // fused quantiser: descriptor of each planned block (plan order)
    const int32_t *tu;             // fused quantiser: caller's index of each planned block
    int n;
    const int8_t *mats;
    int matOff[3][7];
    int matCols[3][7];
    int diagOff[3][4];
};


// Zeroes the coefficients outside the kept Hk x Wk corner of every block of a
// group with V-sized stores (V = int4 for widths >= 4; int2 for 2-wide blocks,
// whose offsets are only 8-byte aligned).  The zero rows Hk..H-1 of a block are
// one contiguous run; the columns right of Wk exist only for widths 32 and 64,
// where W - Wk >= 16.  All counts are powers of two.
template <typename V, typename E>
__device__ __forceinline__ void zero_out(E *Og, int myRel, int G_n, int W, int H, int Wk, int Hk, int lane)
{
    constexpr int g = sizeof(V) / sizeof(E);
    const int tail = (W * (H - Hk)) / g, side = (W - Wk) / g;
    const int lgSide = side ? 31 - __builtin_clz(side) : 0;
    V z;
    memset(&z, 0, sizeof(z));
    for (int b = 0; b < G_n; b++) {
        E *O = Og + __builtin_amdgcn_readlane(myRel, b);
        for (int e = lane; e < tail; e += 64) *reinterpret_cast<V *>(O + Hk * W + g * e) = z;
        for (int e = lane; e < (Hk << lgSide) && side; e += 64) {
            const int j = e >> lgSide;
            *reinterpret_cast<V *>(O + j * W + Wk + g * (e & (side - 1))) = z;
        }
    }
}

// int8 limbs of 8 samples (lo = r & 127, hi = r >> 7), bytes in sample order
__device__ __forceinline__ void limbs8(uint4 v, long &blo, long &bhi)
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t m[4], hh[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        m[q] = w[q] & 0x007F007Fu;
        hh[q] = as_u32(as_v2s(w[q]) >> (v2s){7, 7});
    }
    blo = bytes8_lo16(m);
    bhi = bytes8_lo16(hh);
}

// One workgroup = one wave; it walks groups of same-descriptor blocks.  A group
// issues all its global loads (residual chunks and basis fragments) before its
// first wait and fills the LDS stage, so one memory latency is exposed per
// group of up to 64 blocks.  Both stages put the data in the A operand and the
// basis in B, so D comes out transposed: stage 1 leaves four consecutive rows
// y of one column k per lane (packed straight into the limb planes), stage 2
// four consecutive columns k of one output row j (one 16-byte store).
// QUANT: the quantiser of each block is applied to those four coefficients
// in registers (they are one row of a 4x4 coefficient group, whose other rows
// are the lanes of the same quad), and the levels and dequantised values are
// stored instead of the coefficients; the group maps collect in LDS.
template <int HMAX, bool QUANT>
__global__ void __launch_bounds__(64) k_fdct_mfma(FdctArgs a, const FdctGroup *__restrict__ groups,
                                                   const int32_t *__restrict__ rel)
{
    constexpr int NC = FdctLds<HMAX>::kStage / 512;   // 16-byte chunks per lane
    constexpr int PL = FdctLds<HMAX>::kPlane;
    __shared__ uint4 stg[64 * NC];
    __shared__ uint64_t limb[3 * PL / 8];
    const int16_t *S = reinterpret_cast<const int16_t *>(stg);
    uint8_t *P0 = reinterpret_cast<uint8_t *>(limb), *P1 = P0 + PL, *P2 = P1 + PL;
    const int lane = threadIdx.x;
    const int lr = lane & 15, lq = lane >> 4;
    for (int gi = blockIdx.x; gi < a.n; gi += gridDim.x) {
        const FdctGroup G = groups[gi];
        const int G_n = G.count;
        const int myRel = rel[G.first + min(lane, G_n - 1)];
        uint32_t myQd = 0;                                   // quantiser descriptor of block `lane`
        if constexpr (QUANT) myQd = a.qd[G.first + min(lane, G_n - 1)];
        const uint32_t desc = G.desc;
        const int lgW = desc & 15, lgH = (desc >> 4) & 15;
        const int hor = (desc >> 8) & 15, ver = (desc >> 12) & 15;
        const int W = 1 << lgW, H = 1 << lgH;
        const int Wk = tr_kept(hor, W), Hk = tr_kept(ver, H);
        const int lgWk = 31 - __builtin_clz(Wk);
        const int TS = tmp_stride(H);
        const int s1 = lgW + 1, s2 = lgH + 6;
        // sides up to 8: block-diagonal basis, 16 / side row sets (stage 1) or
        // columns (stage 2) per tile, so no lane of D holds padding
        const bool dW = W <= 8, dH = H <= 8;
        const int8_t *TW = a.mats + (dW ? a.diagOff[hor][lgW] : a.matOff[hor][lgW]);
        const int8_t *TH = a.mats + (dH ? a.diagOff[ver][lgH] : a.matOff[ver][lgH]);
        const int cw = dW ? 32 : a.matCols[hor][lgW], ch = dH ? 32 : a.matCols[ver][lgH];
        const int16_t *Rg = a.resid + G.base;
        int32_t *Og = a.coeff + G.base;

        // ---- loads: basis fragments (B operands of both stages) and the residual ----
        // unused fragments re-read fragment (0, 0): the loads are unconditional,
        // so the waits below count them exactly
        long a1[2][2], a2[2][2];
#pragma unroll
        for (int mi = 0; mi < 2; mi++)
#pragma unroll
            for (int ki = 0; ki < 2; ki++) {
                const bool u1 = 16 * mi < Wk && 32 * ki < W, u2 = 16 * mi < Hk && 32 * ki < H;
                a1[mi][ki] = load8(TW + (u1 ? (16 * mi + lr) * cw + 32 * ki : lr * cw) + 8 * lq);
                a2[mi][ki] = load8(TH + (u2 ? (16 * mi + lr) * ch + 32 * ki : lr * ch) + 8 * lq);
            }
        if (W >= 8) {
            const int lgCpb = lgW + lgH - 3, nch = G_n << lgCpb;
            uint4 c[NC];
            stage_group(c, Rg, myRel, lgCpb, nch, 8, lane);
            stage_store(c, stg, lane);
        } else if (W == 4) {
            const int lgCpb = lgH, nch = G_n << lgCpb;
            uint2 c[2 * NC];
            stage_group(c, Rg, myRel, lgCpb, nch, 4, lane);
            stage_store(c, stg, lane);
        } else {
            const int lgCpb = lgH, nch = G_n << lgCpb;
            uint32_t c[4 * NC];
            stage_group(c, Rg, myRel, lgCpb, nch, 2, lane);
            stage_store(c, stg, lane);
        }
        __syncthreads();

        // ---- stage 1: M = rows (b, y) of the group, N = k < Wk, K = x < W ----
        const int rows1 = G_n * H;
        const int rnd1 = 1 << (s1 - 1);
        const int lgQ1 = dW ? 4 - lgW : 0;                 // row sets per tile
        for (int n0 = 0; n0 < rows1; n0 += 16 << lgQ1) {
            long blo[2], bhi[2];
            if (!dW) {
                const int row = n0 + lr;
                const int16_t *R = S + (row < rows1 ? row << lgW : 0);     // blocks are back to back
#pragma unroll
                for (int ki = 0; ki < 2; ki++) {
                    const int x = 32 * ki + 8 * lq;
                    const uint4 v = x < W ? *reinterpret_cast<const uint4 *>(R + x) : make_uint4(0, 0, 0, 0);
                    limbs8(v, blo[ki], bhi[ki]);
                }
            } else {
                // K slice lq < 2 carries row sets q = lq * 8 / W .. : row n0 + lr + 16 q, W samples each
                uint4 v = make_uint4(0, 0, 0, 0);
                if (lq < 2) {
                    if (W == 8) {
                        const int r0 = n0 + lr + 16 * lq;
                        if (r0 < rows1) v = *reinterpret_cast<const uint4 *>(S + r0 * 8);
                    } else if (W == 4) {
                        const int r0 = n0 + lr + 32 * lq, r1 = r0 + 16;
                        if (r0 < rows1) {
                            const uint2 h = *reinterpret_cast<const uint2 *>(S + r0 * 4);
                            v.x = h.x;
                            v.y = h.y;
                        }
                        if (r1 < rows1) {
                            const uint2 h = *reinterpret_cast<const uint2 *>(S + r1 * 4);
                            v.z = h.x;
                            v.w = h.y;
                        }
                    } else {
                        uint32_t w[4];
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            const int rt = n0 + lr + 16 * (4 * lq + t);
                            w[t] = rt < rows1 ? *reinterpret_cast<const uint32_t *>(S + rt * 2) : 0;
                        }
                        v = make_uint4(w[0], w[1], w[2], w[3]);
                    }
                }
                limbs8(v, blo[0], bhi[0]);
                blo[1] = bhi[1] = 0;
            }
            // this lane's D rows: n0 + 4 lq + 16 q + r, r = 0..3, of column k
            const int q = dW ? lr >> lgW : 0;
            const int orow = n0 + 4 * lq + 16 * q, ob = orow >> lgH, oy = orow & (H - 1);
#pragma unroll
            for (int mi = 0; mi < 2; mi++) {
                if (16 * mi >= Wk) break;
                v4i dlo = {rnd1, rnd1, rnd1, rnd1}, dhi = {0, 0, 0, 0};     // rounding offset rides in the accumulator
#pragma unroll
                for (int ki = 0; ki < 2; ki++) {
                    if (32 * ki >= W) break;
                    dlo = __builtin_amdgcn_mfma_i32_16x16x32_i8(blo[ki], a1[mi][ki], dlo, 0, 0, 0);
                    dhi = __builtin_amdgcn_mfma_i32_16x16x32_i8(bhi[ki], a1[mi][ki], dhi, 0, 0, 0);
                }
                int t[4];
#pragma unroll
                for (int r = 0; r < 4; r++) t[r] = (dlo[r] + (dhi[r] << 7)) >> s1;
                const uint32_t L0 = bytes4(t[0], t[1], t[2], t[3]) & 0x7F7F7F7Fu;
                const uint32_t L1 = bytes4(t[0] >> 7, t[1] >> 7, t[2] >> 7, t[3] >> 7) & 0x7F7F7F7Fu;
                const uint32_t L2 = bytes4(t[0] >> 14, t[1] >> 14, t[2] >> 14, t[3] >> 14);
                const int k = dW ? lr & (W - 1) : 16 * mi + lr;
                if (k < Wk && ob < G_n) {
                    const int o = ((ob << lgWk) + k) * TS + oy;
                    if (H >= 4) {
                        *reinterpret_cast<uint32_t *>(P0 + o) = L0;
                        *reinterpret_cast<uint32_t *>(P1 + o) = L1;
                        *reinterpret_cast<uint32_t *>(P2 + o) = L2;
                    } else {
                        // 2-high blocks: rows r = 0, 1 belong to block ob, rows 2, 3 to block ob + 1
                        *reinterpret_cast<uint16_t *>(P0 + o) = (uint16_t)L0;
                        *reinterpret_cast<uint16_t *>(P1 + o) = (uint16_t)L1;
                        *reinterpret_cast<uint16_t *>(P2 + o) = (uint16_t)L2;
                        if (ob + 1 < G_n) {
                            const int o2 = o + (Wk * TS);
                            *reinterpret_cast<uint16_t *>(P0 + o2) = (uint16_t)(L0 >> 16);
                            *reinterpret_cast<uint16_t *>(P1 + o2) = (uint16_t)(L1 >> 16);
                            *reinterpret_cast<uint16_t *>(P2 + o2) = (uint16_t)(L2 >> 16);
                        }
                    }
                }
            }
        }
        __syncthreads();

        // fused quantiser: the group maps live in the residual stage, which
        // stage 1 has finished reading
        unsigned long long *nzl = reinterpret_cast<unsigned long long *>(stg);
        if constexpr (QUANT) {
            nzl[lane] = 0;
            __syncthreads();
        }

        // ---- stage 2: M = columns (b, k) of the group, N = j < Hk, K = y < H ----
        const int cols2 = G_n * Wk;
        const int rnd2 = 1 << (s2 - 1);
        const int lgQ2 = dH ? 4 - lgH : 0;                 // column sets per tile
        for (int n0 = 0; n0 < cols2; n0 += 16 << lgQ2) {
            long p0[2], p1[2], p2[2];
            if (!dH) {
                const int col = n0 + lr;
                const int lrow = (col < cols2 ? col : 0) * TS;        // limb row of this column
#pragma unroll
                for (int ki = 0; ki < 2; ki++) {
                    const int yb = 32 * ki + 8 * lq;
                    p0[ki] = p1[ki] = p2[ki] = 0;
                    if (yb < H) {
                        p0[ki] = *reinterpret_cast<const long *>(P0 + lrow + yb);
                        p1[ki] = *reinterpret_cast<const long *>(P1 + lrow + yb);
                        p2[ki] = *reinterpret_cast<const long *>(P2 + lrow + yb);
                    }
                }
            } else {
                // K slice lq < 2 carries column sets q = lq * 8 / H .. : limb row n0 + lr + 16 q, H bytes each
                unsigned long long x0 = 0, x1 = 0, x2 = 0;
                if (lq < 2) {
                    if (H == 8) {
                        const int c0 = n0 + lr + 16 * lq;
                        if (c0 < cols2) {
                            x0 = *reinterpret_cast<const unsigned long long *>(P0 + c0 * 8);
                            x1 = *reinterpret_cast<const unsigned long long *>(P1 + c0 * 8);
                            x2 = *reinterpret_cast<const unsigned long long *>(P2 + c0 * 8);
                        }
                    } else if (H == 4) {
#pragma unroll
                        for (int t = 0; t < 2; t++) {
                            const int ct = n0 + lr + 16 * (2 * lq + t);
                            if (ct < cols2) {
                                x0 |= (unsigned long long)*reinterpret_cast<const uint32_t *>(P0 + ct * 8) << (32 * t);
                                x1 |= (unsigned long long)*reinterpret_cast<const uint32_t *>(P1 + ct * 8) << (32 * t);
                                x2 |= (unsigned long long)*reinterpret_cast<const uint32_t *>(P2 + ct * 8) << (32 * t);
                            }
                        }
                    } else {
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            const int ct = n0 + lr + 16 * (4 * lq + t);
                            if (ct < cols2) {
                                x0 |= (unsigned long long)*reinterpret_cast<const uint16_t *>(P0 + ct * 8) << (16 * t);
                                x1 |= (unsigned long long)*reinterpret_cast<const uint16_t *>(P1 + ct * 8) << (16 * t);
                                x2 |= (unsigned long long)*reinterpret_cast<const uint16_t *>(P2 + ct * 8) << (16 * t);
                            }
                        }
                    }
                }
                p0[0] = (long)x0;
                p1[0] = (long)x1;
                p2[0] = (long)x2;
                p0[1] = p1[1] = p2[1] = 0;
            }
            // this lane's D columns: n0 + 4 lq + 16 q + r, r = 0..3, of output row j
            const int q = dH ? lr >> lgH : 0;
            const int ocol = n0 + 4 * lq + 16 * q, obk = ocol >> lgWk, ok0 = ocol & (Wk - 1);
            const int orel = __shfl(myRel, obk < G_n ? obk : 0);
            const int orel2 = __shfl(myRel, obk + 1 < G_n ? obk + 1 : 0);
#pragma unroll
            for (int mi = 0; mi < 2; mi++) {
                if (16 * mi >= Hk) break;
                v4i d0 = {rnd2, rnd2, rnd2, rnd2}, d1 = {0, 0, 0, 0}, d2 = {0, 0, 0, 0};
#pragma unroll
                for (int ki = 0; ki < 2; ki++) {
                    if (32 * ki >= H) break;
                    d0 = __builtin_amdgcn_mfma_i32_16x16x32_i8(p0[ki], a2[mi][ki], d0, 0, 0, 0);
                    d1 = __builtin_amdgcn_mfma_i32_16x16x32_i8(p1[ki], a2[mi][ki], d1, 0, 0, 0);
                    d2 = __builtin_amdgcn_mfma_i32_16x16x32_i8(p2[ki], a2[mi][ki], d2, 0, 0, 0);
                }
                int v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = (d0[r] + (d1[r] << 7) + (d2[r] << 14)) >> s2;
                const int j = dH ? lr & (H - 1) : 16 * mi + lr;
                if constexpr (QUANT) {
                    // Wk >= 4 and H >= 4 here (4x4 groups): v[] is one row of a group
                    const QParams qp = quant_params((uint32_t)__shfl((int)myQd, obk < G_n ? obk : 0));
                    int qv[4], rv[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) qv[r] = qiq24(v[r], qp, rv[r]);
                    const bool live = j < Hk && obk < G_n;
                    if (live) {
                        const int o = orel + j * W + ok0;
                        *reinterpret_cast<uint2 *>(a.q + G.base + o) = make_uint2(pack16(qv[0], qv[1]),
                                                                                  pack16(qv[2], qv[3]));
                        *reinterpret_cast<uint2 *>(a.r + G.base + o) = make_uint2(pack16(rv[0], rv[1]),
                                                                                  pack16(rv[2], rv[3]));
                    }
                    int f = live && (qv[0] | qv[1] | qv[2] | qv[3]) != 0;
                    f |= shfl_xor<1>(f);                    // the quad holds the group's four rows
                    f |= shfl_xor<2>(f);
                    if (live && f && (lr & 3) == 0)
                        atomicOr(&nzl[obk], 1ull << ((j >> 2) * (min(W, 32) >> 2) + (ok0 >> 2)));
                    continue;
                }
                if (j < Hk && obk < G_n) {
                    if (Wk >= 4) {
                        *reinterpret_cast<int4 *>(Og + orel + j * W + ok0) = make_int4(v[0], v[1], v[2], v[3]);
                    } else {
                        // 2-wide blocks: columns r = 0, 1 belong to block obk, 2, 3 to block obk + 1
                        *reinterpret_cast<int2 *>(Og + orel + j * W) = make_int2(v[0], v[1]);
                        if (obk + 1 < G_n) *reinterpret_cast<int2 *>(Og + orel2 + j * W) = make_int2(v[2], v[3]);
                    }
                }
            }
        }
        // zero-out region: rows Hk..H-1 whole, then columns Wk..W-1 of the kept rows
        if constexpr (QUANT) {
            // levels and dequantised values of zero coefficients are zero
            if (Wk < W || Hk < H) {
                zero_out<uint2>(a.q + G.base, myRel, G_n, W, H, Wk, Hk, lane);
                zero_out<uint2>(a.r + G.base, myRel, G_n, W, H, Wk, Hk, lane);
            }
            __syncthreads();
            if (lane < G_n) a.nz[a.tu[G.first + lane]] = nzl[lane];
        } else if (Wk < W || Hk < H) {
            if (W >= 4)
                zero_out<int4>(Og, myRel, G_n, W, H, Wk, Hk, lane);
            else
                zero_out<int2>(Og, myRel, G_n, W, H, Wk, Hk, lane);
        }
        __syncthreads();
    }
}

}  // namespace xg

// ---------------------------------------------------------------------------
// Plan / run
// ---------------------------------------------------------------------------
using namespace xg;

namespace {
struct FdctPlan {
    FdctGroup *dGroups = nullptr;     // groups of blocks up to 32 high first, then the 64-high ones
    int32_t *dRel = nullptr;          // block offsets relative to their group's base
    uint32_t *dQd = nullptr;          // fused quantiser: block descriptors in plan order
    int32_t *dTu = nullptr;           // fused quantiser: caller's index of each block in plan order
    int n = 0, nLow = 0, cap = 0;     // groups
    int relCap = 0, qdCap = 0, tuCap = 0;
    bool valid = false, quant = false;
};
std::mutex g_plan_mu;
std::vector<std::pair<xg_session *, FdctPlan>> g_plans;

FdctPlan *plan_for(xg_session *s, bool create)
{
    for (auto &p : g_plans)
        if (p.first == s) return &p.second;
    if (!create) return nullptr;
    g_plans.push_back({s, FdctPlan()});
    return &g_plans.back().second;
}

// a plan that failed validation leaves no runnable plan behind
int32_t plan_fail(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_plan_mu);
    if (FdctPlan *p = plan_for(s, false)) p->valid = false;
    return XG_ERR_ARG;
}

}  // namespace

// The forward-transform plan; with qdesc, also the quantiser descriptors of a
// fused quantising run (4x4 coefficient groups, sides matching the transform's).
static int32_t fdct_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, const uint32_t *qdesc,
                         int32_t n, const char *fn)
{
    if (!s || n < 0 || (n > 0 && (!offset || !desc))) {
        set_error(XG_ERR_ARG, "%s: bad arguments", fn);
        return s ? plan_fail(s) : XG_ERR_ARG;
    }
    if (!mats_init()) return XG_ERR_HIP;
    static const int8_t mts[6][2] = {{kDCT2, kDCT2}, {kSKIP, kSKIP}, {kDST7, kDST7},
                                     {kDCT8, kDST7}, {kDST7, kDCT8}, {kDCT8, kDCT8}};
    struct Item {
        long long off;
        uint32_t desc;
        int32_t idx;
    };
    std::vector<Item> items((size_t)n);
    for (int i = 0; i < n; i++) {
        const int lgW = desc[i] & 15, lgH = (desc[i] >> 4) & 15, m = (desc[i] >> 8) & 15;
        if (lgW < 1 || lgW > 6 || lgH < 1 || lgH > 6 || m > 5 || m == 1) {
            set_error(XG_ERR_ARG, "%s: block %d has descriptor 0x%x", fn, i, desc[i]);
            return plan_fail(s);
        }
        if (qdesc) {
            const uint32_t q = qdesc[i];
            if ((q & 0xFFF) != (desc[i] & 0xFFF) || lgW < 2 || lgH < 2 || ((q >> 12) & 7) != 2 ||
                ((q >> 15) & 7) != 2 || (q >> 26)) {
                set_error(XG_ERR_ARG, "%s: block %d: quantiser descriptor 0x%x must repeat the sides and mtsIdx "
                                      "of 0x%x, with 4x4 coefficient groups (sides >= 4)", fn, i, q, desc[i]);
                return plan_fail(s);
            }
        }
        const int hor = mts[m][0], ver = mts[m][1];
        if ((hor != kDCT2 && (lgW < 2 || lgW > 5)) || (ver != kDCT2 && (lgH < 2 || lgH > 5)) || offset[i] < 0) {
            set_error(XG_ERR_ARG, "%s: block %d: DST-VII/DCT-VIII need sides 4..32", fn, i);
            return plan_fail(s);
        }
        if (offset[i] % std::min(8, 1 << lgW)) {
            set_error(XG_ERR_ARG, "%s: block %d: offset must be a multiple of min(8, width) samples", fn, i);
            return plan_fail(s);
        }
        items[i].off = offset[i];
        items[i].desc = (uint32_t)(lgW | (lgH << 4) | (hor << 8) | (ver << 12));
        items[i].idx = i;
    }
    // two launches: blocks up to 32 high (small LDS tile), then 64-high blocks;
    // inside each, descriptors with the largest blocks first, blocks of one
    // descriptor in memory order
    std::sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
        const int hx = ((x.desc >> 4) & 15) == 6, hy = ((y.desc >> 4) & 15) == 6;
        if (hx != hy) return hx < hy;
        const int ax = (x.desc & 15) + ((x.desc >> 4) & 15), ay = (y.desc & 15) + ((y.desc >> 4) & 15);
        if (ax != ay) return ax > ay;
        if (x.desc != y.desc) return x.desc < y.desc;
        if (x.off != y.off) return x.off < y.off;
        return x.idx < y.idx;
    });
    std::vector<FdctGroup> groups;
    std::vector<int32_t> rel((size_t)n);
    int nLow = 0;
    for (int i = 0; i < n;) {
        const uint32_t d = items[i].desc;
        const int lgW = d & 15, lgH = (d >> 4) & 15;
        const int W = 1 << lgW, H = 1 << lgH;
        const int cap = lgH == 6 ? group_cap(W, H, tr_kept((d >> 8) & 15, W), FdctLds<64>::kStage, FdctLds<64>::kPlane)
                                 : group_cap(W, H, tr_kept((d >> 8) & 15, W), FdctLds<32>::kStage, FdctLds<32>::kPlane);
        FdctGroup g;
        memset(&g, 0, sizeof(g));
        g.desc = d;
        g.base = items[i].off;
        g.first = i;
        // relative offsets are int32: a group spans less than 2^30 samples
        while (i < n && items[i].desc == d && g.count < cap && items[i].off - g.base < (1LL << 30)) {
            rel[(size_t)i] = (int32_t)(items[i].off - g.base);
            g.count++;
            i++;
        }
        groups.push_back(g);
        if (lgH != 6) nLow = (int)groups.size();
    }
    const int ng = (int)groups.size();
    std::lock_guard<std::mutex> g(g_plan_mu);
    FdctPlan *p = plan_for(s, true);
    p->valid = false;
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    if (!grow(p->dGroups, p->cap, std::max(ng, 1), st) || !grow(p->dRel, p->relCap, std::max(n, 1), st))
        return XG_ERR_NOMEM;
    if (ng && !XG_HIP(hipMemcpyAsync(p->dGroups, groups.data(), sizeof(FdctGroup) * (size_t)ng, hipMemcpyHostToDevice,
                                     st)))
        return XG_ERR_HIP;
    if (n && !XG_HIP(hipMemcpyAsync(p->dRel, rel.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st)))
        return XG_ERR_HIP;
    std::vector<uint32_t> qd;
    std::vector<int32_t> tu;
    if (qdesc) {
        qd.resize((size_t)n);
        tu.resize((size_t)n);
        for (int i = 0; i < n; i++) {
            qd[(size_t)i] = qdesc[items[i].idx];
            tu[(size_t)i] = items[i].idx;
        }
        if (!grow(p->dQd, p->qdCap, std::max(n, 1), st) || !grow(p->dTu, p->tuCap, std::max(n, 1), st))
            return XG_ERR_NOMEM;
        if (n && (!XG_HIP(hipMemcpyAsync(p->dQd, qd.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice,
                                         st)) ||
                  !XG_HIP(hipMemcpyAsync(p->dTu, tu.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st))))
            return XG_ERR_HIP;
    }
    if (!XG_HIP(hipStreamSynchronize(st))) return XG_ERR_HIP;
    p->n = ng;
    p->nLow = nLow;
    p->quant = qdesc != nullptr;
    p->valid = true;
    return XG_OK;
}

extern "C" {

int32_t xg_fdct_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, int32_t n)
{
    return fdct_plan(s, offset, desc, nullptr, n, "xg_fdct_plan");
}

int32_t xg_fdct_quant_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, const uint32_t *qdesc,
                           int32_t n)
{
    if (n > 0 && !qdesc) {
        set_error(XG_ERR_ARG, "xg_fdct_quant_plan: bad arguments");
        return s ? plan_fail(s) : XG_ERR_ARG;
    }
    static const uint32_t kNone = 0;     // an empty plan is still a quantising one
    return fdct_plan(s, offset, desc, qdesc ? qdesc : &kNone, n, "xg_fdct_quant_plan");
}

int32_t xg_fdct_run(xg_session *s, const int16_t *resid, xg_coeff *coeff)
{
    if (!s || !resid || !coeff) {
        set_error(XG_ERR_ARG, "xg_fdct_run: bad arguments");
        return XG_ERR_ARG;
    }
    FdctPlan *p;
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        p = plan_for(s, false);
    }
    if (!p || !p->valid) {
        set_error(XG_ERR_STATE, "xg_fdct_run: no plan (call xg_fdct_plan first; a failed plan leaves none)");
        return XG_ERR_STATE;
    }
    if (p->n == 0) return XG_OK;
    FdctArgs a;
    memset(&a, 0, sizeof(a));
    a.resid = resid;
    a.coeff = coeff;
    a.mats = g_mats;
    memcpy(a.matOff, g_mi.off, sizeof(a.matOff));
    memcpy(a.matCols, g_mi.cols, sizeof(a.matCols));
    memcpy(a.diagOff, g_mi.diag, sizeof(a.diagOff));
    // one resident wave per workgroup slot the chip holds (registers and LDS bound it)
    static const int slots[2] = {resident_slots((const void *)k_fdct_mfma<32, false>),
                                 resident_slots((const void *)k_fdct_mfma<64, false>)};
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    int ev[2];
    const bool timed = session_timer_begin(s, 2, ev);
    if (p->nLow > 0) {
        FdctArgs lo = a;
        lo.n = p->nLow;
        k_fdct_mfma<32, false><<<std::min(lo.n, slots[0]), 64, 0, st>>>(lo, p->dGroups, p->dRel);
    }
    if (p->n > p->nLow) {
        FdctArgs hi = a;
        hi.n = p->n - p->nLow;
        k_fdct_mfma<64, false><<<std::min(hi.n, slots[1]), 64, 0, st>>>(hi, p->dGroups + p->nLow, p->dRel);
    }
    if (timed) session_timer_end(s, 2, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

int32_t xg_fdct_quant_run(xg_session *s, const int16_t *resid, int16_t *qCoeff, int16_t *rCoeff, uint64_t *nzMap)
{
    if (!s || !resid || !qCoeff || !rCoeff || !nzMap || ((uintptr_t)qCoeff & 7) || ((uintptr_t)rCoeff & 7) ||
        ((uintptr_t)nzMap & 7)) {
        set_error(XG_ERR_ARG, "xg_fdct_quant_run: bad arguments (outputs 8-byte aligned)");
        return XG_ERR_ARG;
    }
    FdctPlan *p;
    {
        std::lock_guard<std::mutex> g(g_plan_mu);
        p = plan_for(s, false);
    }
    if (!p || !p->valid || !p->quant) {
        set_error(XG_ERR_STATE, "xg_fdct_quant_run: no quantising plan (call xg_fdct_quant_plan first)");
        return XG_ERR_STATE;
    }
    if (p->n == 0) return XG_OK;
    FdctArgs a;
    memset(&a, 0, sizeof(a));
    a.resid = resid;
    a.q = qCoeff;
    a.r = rCoeff;
    a.nz = reinterpret_cast<unsigned long long *>(nzMap);
    a.qd = p->dQd;
    a.tu = p->dTu;
    a.mats = g_mats;
    memcpy(a.matOff, g_mi.off, sizeof(a.matOff));
    memcpy(a.matCols, g_mi.cols, sizeof(a.matCols));
    memcpy(a.diagOff, g_mi.diag, sizeof(a.diagOff));
    static const int slots[2] = {resident_slots((const void *)k_fdct_mfma<32, true>),
                                 resident_slots((const void *)k_fdct_mfma<64, true>)};
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    int ev[2];
    const bool timed = session_timer_begin(s, 2, ev);
    if (p->nLow > 0) {
        FdctArgs lo = a;
        lo.n = p->nLow;
        k_fdct_mfma<32, true><<<std::min(lo.n, slots[0]), 64, 0, st>>>(lo, p->dGroups, p->dRel);
    }
    if (p->n > p->nLow) {
        FdctArgs hi = a;
        hi.n = p->n - p->nLow;
        k_fdct_mfma<64, true><<<std::min(hi.n, slots[1]), 64, 0, st>>>(hi, p->dGroups + p->nLow, p->dRel);
    }
    if (timed) session_timer_end(s, 2, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

}  // extern "C"

void xg::fdct_release(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_plan_mu);
    for (size_t i = 0; i < g_plans.size(); i++)
        if (g_plans[i].first == s) {
            if (g_plans[i].second.dGroups) (void)hipFree(g_plans[i].second.dGroups);
            if (g_plans[i].second.dRel) (void)hipFree(g_plans[i].second.dRel);
            if (g_plans[i].second.dQd) (void)hipFree(g_plans[i].second.dQd);
            if (g_plans[i].second.dTu) (void)hipFree(g_plans[i].second.dTu);
            g_plans.erase(g_plans.begin() + (long)i);
            return;
        }
}

// xg_search.hip -- integer and fractional motion search.
//
//  * k_full_search: exhaustive integer search of one block per workgroup; the
//    block's reference window (every candidate position plus the block
//    extent) and the current block are staged once in LDS; each thread costs
//    a strided subset of candidates; the argmin is a 64-bit min over
//    (cost << 32 | raster index), so ties resolve to the earliest candidate
//    in raster order (dy outer, dx inner) whatever the thread schedule.
//    Cost = (SAD << 8) + (EG(4*dx - pmx) + EG(4*dy - pmy)) * lambda
//    (XIN_COST_FRACTION 8, common/include/h26x_definition.h:88; MVD in
//    quarter pel as CALC_MVD_BIT is used at h266_sub_pel_refine.c:155-156).
//  * Xg26xMotionEstDia: start candidates then small-diamond descent, one
//    workgroup, SADs reduced across its 256 threads.
//  * Xg266FullSubPelMe: the half-pel and quarter-pel rings of
//    h266_encoder/source/h266_sub_pel_refine.c:81-212 with the 8-tap luma
//    interpolation (filterIndex 0) and SAD.
#include <algorithm>
#include <string.h>

#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

__device__ __forceinline__ unsigned long long block_min_u64(unsigned long long v, unsigned long long *red)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const unsigned long long o = (unsigned long long)__shfl_xor((long long)v, m, 64);
        v = o < v ? o : v;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    unsigned long long r = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) r = red[i] < r ? red[i] : r;
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t *red)
{
    v = (uint32_t)group_sum<64>((int)v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    uint32_t r = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += red[i];
    __syncthreads();
    return r;
}

// One block's full search over window [x0,x1] x [y0,y1] (integer MVs).
// win: reference samples from (x0, y0) relative to the block, stride ws;
// cur: the block, stride bw.  Both in LDS.
__device__ unsigned long long full_search_core(const uint16_t *win, int ws, const uint16_t *cur, int bw, int bh,
                                               int x0, int x1, int y0, int y1, uint32_t lambda, int pmx, int pmy,
                                               unsigned long long *red)
{
    const int nx = x1 - x0 + 1, n = nx * (y1 - y0 + 1);
    unsigned long long best = ~0ull;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int ox = i % nx, oy = i / nx;
        const uint16_t *w = win + oy * ws + ox;
        uint32_t sad = 0;
        for (int r = 0; r < bh; r++) {
            const uint16_t *wr = w + r * ws, *cr = cur + r * bw;
            for (int c = 0; c < bw; c++) sad += (uint32_t)abs((int)cr[c] - (int)wr[c]);
        }
        const int mx = x0 + ox, my = y0 + oy;
        const uint32_t cost = (sad << 8) + (eg_bits(4 * mx - pmx) + eg_bits(4 * my - pmy)) * lambda;
        const unsigned long long key = ((unsigned long long)cost << 32) | (uint32_t)i;
        best = key < best ? key : best;
    }
    return block_min_u64(best, red);
}

struct FsFrameArgs {
    const uint16_t *cur; intptr_t cs;
    const uint16_t *ref; intptr_t rs;
    int fw, fh, bw, bh, range, ctu, nbx;
    uint32_t lambda;
    int16_t *bestMv; uint32_t *bestCost; uint32_t *bestSad;
};

// Frame form: block b of the bw x bh grid, predictor (0,0), window
// [-range, range]^2 clipped to the CU MV window (h266_enc_init.c:1462-1465).
__global__ void __launch_bounds__(256) k_full_search_frame(FsFrameArgs a)
{
    extern __shared__ uint16_t smem[];
    __shared__ unsigned long long red[4];
    const int b = blockIdx.x;
    const int x = (b % a.nbx) * a.bw, y = (b / a.nbx) * a.bh;
    const int mnx = (int16_t)(-x - a.ctu - 8), mny = (int16_t)(-y - a.ctu - 8);
    const int mxx = (int16_t)(a.fw + (a.ctu - a.bw + 8) - x), mxy = (int16_t)(a.fh + (a.ctu - a.bh + 8) - y);
    const int x0 = max(-a.range, mnx), x1 = min(a.range, mxx);
    const int y0 = max(-a.range, mny), y1 = min(a.range, mxy);
    const int ws = x1 - x0 + a.bw, wh = y1 - y0 + a.bh;
    uint16_t *win = smem, *cur = smem + ws * wh;
    const uint16_t *rp = a.ref + (intptr_t)(y + y0) * a.rs + (x + x0);
    for (int i = threadIdx.x; i < ws * wh; i += blockDim.x) win[i] = rp[(intptr_t)(i / ws) * a.rs + (i % ws)];
    const uint16_t *cp = a.cur + (intptr_t)y * a.cs + x;
    for (int i = threadIdx.x; i < a.bw * a.bh; i += blockDim.x) cur[i] = cp[(intptr_t)(i / a.bw) * a.cs + (i % a.bw)];
    __syncthreads();
    const unsigned long long k = full_search_core(win, ws, cur, a.bw, a.bh, x0, x1, y0, y1, a.lambda, 0, 0, red);
    if (threadIdx.x == 0) {
        const int i = (int)(uint32_t)k, nx = x1 - x0 + 1;
        const int mx = x0 + i % nx, my = y0 + i / nx;
        a.bestMv[2 * b] = (int16_t)(4 * mx);
        a.bestMv[2 * b + 1] = (int16_t)(4 * my);
        a.bestCost[b] = (uint32_t)(k >> 32);
        uint32_t sad = 0;
        const uint16_t *w = win + (my - y0) * ws + (mx - x0);
        for (int r = 0; r < a.bh; r++)
            for (int c = 0; c < a.bw; c++) sad += (uint32_t)abs((int)cur[r * a.bw + c] - (int)w[r * ws + c]);
        a.bestSad[b] = sad;
    }
}

// ---------------------------------------------------------------------------
// Packed integer search core (block widths 4..64, powers of two)
//
// Each lane costs K horizontally adjacent candidates (Kq .. Kq + K - 1) of one
// search row dyi: a window row is read as BW/2 + K/2 dwords D (sample pairs
// from an even column), F = D realigned by one sample, and v_sad_u16 takes two
// absolute differences per instruction against the packed current row C.
// Window rows are ws samples (ws even, >= nx + BW + K - 1, so the reads of the
// last item stay inside the row); the block is bh rows of BW/2 dwords at a
// 16-byte boundary.  Cost = (SAD << 8) + (EG(4 mx - pmx) + EG(4 my - pmy)) *
// lambda; key = cost << 32 | raster index, as in full_search_core.
// ---------------------------------------------------------------------------

// rows x cols samples from src (row pitch rs) into win (row pitch ws): a wave
// per row, lanes along it, four rows of loads in flight per lane.
__device__ __forceinline__ void stage_rows(uint16_t *win, int ws, const uint16_t *src, intptr_t rs, int rows, int cols)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int r0 = wave; r0 < rows; r0 += 4 * nw) {
        for (int c0 = lane; c0 < cols; c0 += 64) {
            uint16_t v[4];
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = src[(intptr_t)min(r0 + k * nw, rows - 1) * rs + c0];
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (r0 + k * nw < rows) win[(r0 + k * nw) * ws + c0] = v[k];
        }
    }
}

// The block's rows as dword pairs (x even, plane dword aligned).
__device__ __forceinline__ void stage_cur(uint32_t *curL, const uint16_t *cp, intptr_t cs, int bw, int bh)
{
    const int nd = bw >> 1;
    for (int i = threadIdx.x; i < bh * nd; i += blockDim.x)
        curL[i] = *reinterpret_cast<const uint32_t *>(cp + (intptr_t)(i / nd) * cs + 2 * (i % nd));
}

template <int BW, int K = 4>
__device__ __forceinline__ unsigned long long packed_items(const uint32_t *w32, int ws, const uint32_t *curL, int bh,
                                                           int x0, int x1, int y0, int nx, int ny, uint32_t lambda,
                                                           int pmx, int pmy)
{
    constexpr int ND = BW / 2;                      // dword pairs per block row
    constexpr int KD = K / 2;                       // dwords the lane's candidates span beyond the block row
    const int nq = (nx + K - 1) / K, items = ny * nq;
    unsigned long long best = ~0ull;
    for (int it = threadIdx.x; it < items; it += blockDim.x) {
        const int dyi = it / nq, q = it - dyi * nq;
        const uint32_t *rp = w32 + ((dyi * ws) >> 1) + KD * q;
        uint32_t acc[K];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = 0;
#pragma unroll 2
        for (int i = 0; i < bh; i++) {
            uint32_t D[ND + KD], F[ND + KD - 1], C[ND];
#pragma unroll
            for (int j = 0; j < ND + KD; j++) D[j] = rp[i * (ws >> 1) + j];
            if constexpr (ND >= 4) {
                const uint4 *c4 = reinterpret_cast<const uint4 *>(curL + i * ND);
#pragma unroll
                for (int j = 0; j < ND / 4; j++) {
                    const uint4 v = c4[j];
                    C[4 * j] = v.x;
                    C[4 * j + 1] = v.y;
                    C[4 * j + 2] = v.z;
                    C[4 * j + 3] = v.w;
                }
            } else {
#pragma unroll
                for (int j = 0; j < ND; j++) C[j] = curL[i * ND + j];
            }
#pragma unroll
            for (int j = 0; j < ND + KD - 1; j++) F[j] = align16(D[j + 1], D[j], 16);    // pairs from odd samples
#pragma unroll
            for (int j = 0; j < ND; j++)
#pragma unroll
                for (int m = 0; m < KD; m++) {
                    acc[2 * m] = __builtin_amdgcn_sad_u16(D[j + m], C[j], acc[2 * m]);          // candidate K q + 2m
                    acc[2 * m + 1] = __builtin_amdgcn_sad_u16(F[j + m], C[j], acc[2 * m + 1]);  // K q + 2m + 1
                }
        }
        const int my = y0 + dyi;
        const uint32_t by = eg_bits(4 * my - pmy);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int mx = x0 + K * q + k;
            if (mx <= x1) {
                const uint32_t cost = (acc[k] << 8) + (eg_bits(4 * mx - pmx) + by) * lambda;
                const unsigned long long key = ((unsigned long long)cost << 32) | (uint32_t)(dyi * nx + K * q + k);
                best = key < best ? key : best;
            }
        }
    }
    return best;
}

// Minimum over the workgroup (up to 16 waves); every thread gets it.
__device__ __forceinline__ unsigned long long block_min_u64_16(unsigned long long v, unsigned long long *red)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const unsigned long long o = (unsigned long long)__shfl_xor((long long)v, m, 64);
        v = o < v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long r = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) r = red[i] < r ? red[i] : r;
    __syncthreads();
    return r;
}

// Window rows needed by the packed core for nx candidates per row.
__host__ __device__ inline int packed_ws(int nx, int bw, int k = 4) { return (nx + bw + k + 3) & ~1; }

// Candidates per lane of the frame form.  Eight (each window-row read serving
// twice the SADs) measured slightly slower than four at +/-16 and +/-32.
constexpr int kFrameK = 4;

// Threads for a packed search over an n x n candidate window, k candidates
// per lane: the items (search row x run of k candidates) spread evenly over
// the fewest passes a workgroup of at most 1024 threads needs, so no pass runs
// nearly empty (+/-32, k = 4: 1105 items, two passes of 576 threads rather
// than 1024 + 81).
static inline int packed_threads(int n, int k = 4)
{
    const int items = n * ((n + k - 1) / k);
    const int passes = (items + 1023) / 1024;
    return std::min(1024, std::max(64, ((items + passes - 1) / passes + 63) & ~63));
}

// Packed frame form for block widths 8, 16 and 32; same cost, window and tie
// order as k_full_search_frame.
template <int BW>
__global__ void __launch_bounds__(1024) k_full_search_frame_p(FsFrameArgs a, int ws)
{
    extern __shared__ uint32_t smem32[];
    __shared__ unsigned long long red[16];
    const int b = blockIdx.x;
    const int x = (b % a.nbx) * BW, y = (b / a.nbx) * a.bh;
    const int mnx = (int16_t)(-x - a.ctu - 8), mny = (int16_t)(-y - a.ctu - 8);
    const int mxx = (int16_t)(a.fw + (a.ctu - BW + 8) - x), mxy = (int16_t)(a.fh + (a.ctu - a.bh + 8) - y);
    const int x0 = max(-a.range, mnx), x1 = min(a.range, mxx);
    const int y0 = max(-a.range, mny), y1 = min(a.range, mxy);
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1;
    const int rows = ny + a.bh - 1;
    uint16_t *win = reinterpret_cast<uint16_t *>(smem32);
    uint32_t *curL = smem32 + (((rows * (ws >> 1)) + 3) & ~3);    // 16-byte aligned
    stage_rows(win, ws, a.ref + (intptr_t)(y + y0) * a.rs + (x + x0), a.rs, rows, nx + BW - 1);
    stage_cur(curL, a.cur + (intptr_t)y * a.cs + x, a.cs, BW, a.bh);
    __syncthreads();
    unsigned long long best = packed_items<BW, kFrameK>(smem32, ws, curL, a.bh, x0, x1, y0, nx, ny, a.lambda, 0, 0);
    best = block_min_u64_16(best, red);
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const int idx = (int)(uint32_t)best;
        const int mx = x0 + idx % nx, my = y0 + idx / nx;
        // the winner's SAD, recomputed by one wave
        uint32_t sad = 0;
        const uint16_t *wr = win + (my - y0) * ws + (mx - x0);
        const uint16_t *cr = reinterpret_cast<const uint16_t *>(curL);
        for (int e = lane; e < a.bh * BW; e += 64) {
            const int r = e / BW, c = e % BW;
            sad += (uint32_t)abs((int)cr[e] - (int)wr[r * ws + c]);
        }
        sad = (uint32_t)group_sum<64>((int)sad);
        if (lane == 0) {
            a.bestMv[2 * b] = (int16_t)(4 * mx);
            a.bestMv[2 * b + 1] = (int16_t)(4 * my);
            a.bestCost[b] = (uint32_t)(best >> 32);
            a.bestSad[b] = sad;
        }
    }
}

// Per-call form: packed window and block already in global memory.
__global__ void __launch_bounds__(256) k_full_search_one(const uint16_t *gwin, int ws, int wh, const uint16_t *gcur,
                                                         int bw, int bh, int x0, int x1, int y0, int y1,
                                                         uint32_t lambda, int pmx, int pmy, uint32_t *out)
{
    extern __shared__ uint16_t smem[];
    __shared__ unsigned long long red[4];
    uint16_t *win = smem, *cur = smem + ws * wh;
    for (int i = threadIdx.x; i < ws * wh; i += blockDim.x) win[i] = gwin[i];
    for (int i = threadIdx.x; i < bw * bh; i += blockDim.x) cur[i] = gcur[i];
    __syncthreads();
    const unsigned long long k = full_search_core(win, ws, cur, bw, bh, x0, x1, y0, y1, lambda, pmx, pmy, red);
    if (threadIdx.x == 0) {
        const int i = (int)(uint32_t)k, nx = x1 - x0 + 1;
        const int mx = x0 + i % nx, my = y0 + i / nx;
        uint32_t sad = 0;
        const uint16_t *w = win + (my - y0) * ws + (mx - x0);
        for (int r = 0; r < bh; r++)
            for (int c = 0; c < bw; c++) sad += (uint32_t)abs((int)cur[r * bw + c] - (int)w[r * ws + c]);
        out[0] = (uint32_t)mx;
        out[1] = (uint32_t)my;
        out[2] = (uint32_t)(k >> 32);
        out[3] = sad;
    }
}

// SAD of the block at integer offset (mx, my) inside a packed region whose
// origin is (rx0, ry0) relative to the co-located position; block-wide sum.
__device__ uint32_t region_sad(const uint16_t *reg, int rstride, int rx0, int ry0, const uint16_t *cur, int bw,
                               int bh, int mx, int my, uint32_t *red)
{
    const uint16_t *w = reg + (my - ry0) * rstride + (mx - rx0);
    uint32_t s = 0;
    for (int i = threadIdx.x; i < bw * bh; i += blockDim.x) {
        const int r = i / bw, c = i % bw;
        s += (uint32_t)abs((int)cur[i] - (int)w[r * rstride + c]);
    }
    return block_sum_u32(s, red);
}

// Diamond search in one workgroup; the region covers every position the
// search can read.  out: bestX, bestY (integer pel), cost, sad.
__global__ void __launch_bounds__(256) k_dia_one(const uint16_t *reg, int rstride, int rx0, int ry0,
                                                 const uint16_t *cur, int bw, int bh, const int16_t *cand,
                                                 int ncand, int mnx, int mny, int mxx, int mxy, int cx, int cy,
                                                 int sw, int sh, uint32_t lambda, int pmx, int pmy, uint32_t *out)
{
    __shared__ uint32_t red[4];
    uint32_t best = 0xFFFFFFFFu, bestSad = 0xFFFFFFFFu;
    int bx = 0, by = 0;
    for (int i = 0; i < ncand; i++) {
        const int mx = min(max((int)cand[2 * i], mnx), mxx), my = min(max((int)cand[2 * i + 1], mny), mxy);
        const uint32_t sad = region_sad(reg, rstride, rx0, ry0, cur, bw, bh, mx, my, red);
        const uint32_t c = (sad << 8) + (eg_bits(4 * mx - pmx) + eg_bits(4 * my - pmy)) * lambda;
        if (c < best) { best = c; bestSad = sad; bx = mx; by = my; }
    }
    const int dxs[4] = {-1, 1, 0, 0}, dys[4] = {0, 0, -1, 1};
    for (int it = 0; it < sw + sh; it++) {
        int nx = bx, ny = by;
        uint32_t nb = best, ns = bestSad;
        for (int k = 0; k < 4; k++) {
            const int mx = bx + dxs[k], my = by + dys[k];
            if (mx < mnx || mx > mxx || my < mny || my > mxy) continue;
            if (mx < cx - sw || mx > cx + sw || my < cy - sh || my > cy + sh) continue;
            const uint32_t sad = region_sad(reg, rstride, rx0, ry0, cur, bw, bh, mx, my, red);
            const uint32_t c = (sad << 8) + (eg_bits(4 * mx - pmx) + eg_bits(4 * my - pmy)) * lambda;
            if (c < nb) { nb = c; ns = sad; nx = mx; ny = my; }
        }
        if (nx == bx && ny == by) break;
        bx = nx; by = ny; best = nb; bestSad = ns;
    }
    if (threadIdx.x == 0) {
        out[0] = (uint32_t)bx;
        out[1] = (uint32_t)by;
        out[2] = best;
        out[3] = bestSad;
    }
}

// Half-pel then quarter-pel ring around (qx, qy) (quarter pel).  reg: luma
// samples whose first row/column is (rx0, ry0) relative to the co-located
// block position (the host stages integer positions -1..+1 plus 8-tap reach).
__global__ void __launch_bounds__(256) k_subpel_one(const uint16_t *reg, int rstride, int rx0, int ry0,
                                                    const uint16_t *cur, int bw, int bh, int qx, int qy,
                                                    uint32_t bestCost, int pmx, int pmy, uint32_t lambda,
                                                    uint32_t *out)
{
    __shared__ uint32_t red[4];
    const int ring_x[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, ring_y[8] = {0, 0, -1, 1, -1, -1, 1, 1};
    int mx = qx, my = qy;
    uint32_t best = bestCost;
    for (int pass = 0; pass < 2; pass++) {
        const int step = pass == 0 ? 2 : 1;
        const int cx = mx, cy = my;
        for (int k = 0; k < 8; k++) {
            const int tx = cx + ring_x[k] * step, ty = cy + ring_y[k] * step;
            const int m16x = tx * 4, m16y = ty * 4;
            const int8_t *ch = c_luma[0][m16x & 15], *cv = c_luma[0][m16y & 15];
            const int ix = (m16x >> 4) - 3 - rx0, iy = (m16y >> 4) - 3 - ry0;
            uint32_t s = 0;
            for (int i = threadIdx.x; i < bw * bh; i += blockDim.x) {
                const int r = i / bw, c = i % bw;
                const uint16_t *p = reg + (iy + r) * rstride + ix + c;
                int acc = 0;
#pragma unroll
                for (int j = 0; j < 8; j++) acc += (int)cv[j] * hpass<8>(p + j * rstride, ch);
                s += (uint32_t)abs((int)cur[i] - vfinal_pix(acc));
            }
            const uint32_t sad = block_sum_u32(s, red);
            const uint32_t cost = (sad << 8) + (eg_bits(tx - pmx) + eg_bits(ty - pmy)) * lambda;
            if (cost < best) { best = cost; mx = tx; my = ty; }
        }
    }
    if (threadIdx.x == 0) {
        out[0] = (uint32_t)mx;
        out[1] = (uint32_t)my;
        out[2] = best;
    }
}

static inline int clampi_h(int v, int lo, int hi) { return std::min(std::max(v, lo), hi); }

// Integer search of one block (its workgroup): the window +/-range around the
// rounded predictor, clipped to the CU MV window, staged in LDS with the block
// (packed core; ws: window row pitch for the widest block).  Every thread
// returns the winner as the block's base MV (1/16 pel) in B.mvx / B.mvy.
__device__ __forceinline__ void search_block(const xg_picture *pics, xg_block &B, int range, uint32_t lambda, int ctu,
                                             int ws)
{
    extern __shared__ uint32_t smem32[];
    __shared__ unsigned long long red[16];
    const xg_picture &P = pics[B.pic], &R = pics[B.ref];
    const int fw = P.plane[0].width, fh = P.plane[0].height;
    const int mnx = (int16_t)(-B.x - ctu - 8), mny = (int16_t)(-B.y - ctu - 8);
    const int mxx = (int16_t)(fw + (ctu - B.w + 8) - B.x), mxy = (int16_t)(fh + (ctu - B.h + 8) - B.y);
    // predictor: 1/16 pel -> quarter pel (>> 2) -> integer rounded as Xin266FillMeCand does
    // ((v + 2) >> 2, h266_analyze_inter_cu.c:271-272), kept inside the MV window
    const int cx = min(max(((B.mvpx >> 2) + 2) >> 2, mnx), mxx), cy = min(max(((B.mvpy >> 2) + 2) >> 2, mny), mxy);
    const int x0 = max(cx - range, mnx), x1 = min(cx + range, mxx);
    const int y0 = max(cy - range, mny), y1 = min(cy + range, mxy);
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, rows = ny + B.h - 1;
    uint16_t *win = reinterpret_cast<uint16_t *>(smem32);
    uint32_t *curL = smem32 + (((rows * (ws >> 1)) + 3) & ~3);    // 16-byte aligned
    const intptr_t rs = R.plane[0].stride, cs = P.plane[0].stride;
    stage_rows(win, ws, R.plane[0].data + (intptr_t)(B.y + y0) * rs + (B.x + x0), rs, rows, nx + B.w - 1);
    stage_cur(curL, P.plane[0].data + (intptr_t)B.y * cs + B.x, cs, B.w, B.h);
    __syncthreads();
    const int pmx = B.mvpx >> 2, pmy = B.mvpy >> 2;
    unsigned long long k;
    switch (B.w) {
    case 4: k = packed_items<4>(smem32, ws, curL, B.h, x0, x1, y0, nx, ny, lambda, pmx, pmy); break;
    case 8: k = packed_items<8>(smem32, ws, curL, B.h, x0, x1, y0, nx, ny, lambda, pmx, pmy); break;
    case 16: k = packed_items<16>(smem32, ws, curL, B.h, x0, x1, y0, nx, ny, lambda, pmx, pmy); break;
    case 32: k = packed_items<32>(smem32, ws, curL, B.h, x0, x1, y0, nx, ny, lambda, pmx, pmy); break;
    default: k = packed_items<64>(smem32, ws, curL, B.h, x0, x1, y0, nx, ny, lambda, pmx, pmy); break;
    }
    k = block_min_u64_16(k, red);
    const int i = (int)(uint32_t)k;
    B.mvx = 16 * (x0 + i % nx);
    B.mvy = 16 * (y0 + i / nx);
}

// Integer search of every session block around its predictor; the result is
// written back as the block's base MV.
__global__ void __launch_bounds__(1024) k_search_blocks(const xg_picture *pics, xg_block *blocks, int range,
                                                        uint32_t lambda, int ctu, int ws)
{
    const int b = blockIdx.x;
    xg_block B = blocks[b];
    search_block(pics, B, range, lambda, ctu, ws);
    if (threadIdx.x == 0) {
        blocks[b].mvx = B.mvx;
        blocks[b].mvy = B.mvy;
    }
}

}  // namespace xg

using namespace xg;

extern "C" {

void Xg26xMotionEstFull(xg_me_struct *me)
{
    if (!me || !me->input || !me->ref || me->width == 0 || me->height == 0 || me->width > 128 || me->height > 128) {
        set_error(XG_ERR_ARG, "Xg26xMotionEstFull: bad arguments");
        return;
    }
    const int bw = (int)me->width, bh = (int)me->height;
    const int cx = (me->predMv.mvX + 2) >> 2, cy = (me->predMv.mvY + 2) >> 2;
    const int x0 = std::max(cx - me->searchWidth, (int)me->minMv.mvX), x1 = std::min(cx + me->searchWidth, (int)me->maxMv.mvX);
    const int y0 = std::max(cy - me->searchHeight, (int)me->minMv.mvY), y1 = std::min(cy + me->searchHeight, (int)me->maxMv.mvY);
    if (x1 < x0 || y1 < y0) {
        set_error(XG_ERR_ARG, "Xg26xMotionEstFull: empty search window");
        return;
    }
    const int ws = x1 - x0 + bw, wh = y1 - y0 + bh;
    const size_t lds = ((size_t)ws * wh + (size_t)bw * bh) * 2;
    if (lds > 60 * 1024) {
        set_error(XG_ERR_ARG, "Xg26xMotionEstFull: window %dx%d exceeds the LDS budget", ws, wh);
        return;
    }
    const size_t wB = align_up((size_t)ws * wh * 2, 256), cB = align_up((size_t)bw * bh * 2, 256);
    CallCtx *c = call_ctx(wB + cB + 256, wB + cB + 256);
    if (!c) return;
    pack2d(c->host, me->ref + (intptr_t)y0 * me->refStride + x0, me->refStride, ws, wh, 2);
    pack2d(c->host + wB, me->input, me->inputStride, bw, bh, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, wB + cB, hipMemcpyHostToDevice, c->stream))) return;
    uint32_t *dout = (uint32_t *)(c->dev + wB + cB);
    k_full_search_one<<<1, 256, lds, c->stream>>>((uint16_t *)c->dev, ws, wh, (uint16_t *)(c->dev + wB), bw, bh,
                                                  x0, x1, y0, y1, me->sadLambda, me->predMv.mvX,
                                                  me->predMv.mvY, dout);
    if (!XG_HIP(hipGetLastError())) return;
    uint32_t *hout = (uint32_t *)(c->host + wB + cB);
    if (!XG_HIP(hipMemcpyAsync(hout, dout, 16, hipMemcpyDeviceToHost, c->stream))) return;
    if (!XG_HIP(hipStreamSynchronize(c->stream))) return;
    me->bestMv.mvX = (int16_t)(4 * (int)hout[0]);
    me->bestMv.mvY = (int16_t)(4 * (int)hout[1]);
    me->bestCost = hout[2];
    me->bestSad = hout[3];
}

void Xg26xMotionEstDia(xg_me_struct *me)
{
    if (!me || !me->input || !me->ref || me->width == 0 || me->height == 0 || me->width > 128 ||
        me->height > 128 || me->candNum == 0 || me->candNum > 12) {
        set_error(XG_ERR_ARG, "Xg26xMotionEstDia: bad arguments");
        return;
    }
    const int bw = (int)me->width, bh = (int)me->height;
    const int cx = (me->predMv.mvX + 2) >> 2, cy = (me->predMv.mvY + 2) >> 2;
    const int mnx = me->minMv.mvX, mny = me->minMv.mvY, mxx = me->maxMv.mvX, mxy = me->maxMv.mvY;
    // bounding box of every position the search may read
    int x0 = std::max(cx - me->searchWidth, mnx), x1 = std::min(cx + me->searchWidth, mxx);
    int y0 = std::max(cy - me->searchHeight, mny), y1 = std::min(cy + me->searchHeight, mxy);
    int16_t cand[24];
    for (uint32_t i = 0; i < me->candNum; i++) {
        const int mx = clampi_h(me->meCand[i].mvX, mnx, mxx), my = clampi_h(me->meCand[i].mvY, mny, mxy);
        cand[2 * i] = me->meCand[i].mvX;
        cand[2 * i + 1] = me->meCand[i].mvY;
        x0 = std::min(x0, mx); x1 = std::max(x1, mx);
        y0 = std::min(y0, my); y1 = std::max(y1, my);
        // a descent step can leave the window only from a candidate inside it
    }
    const int rw = x1 - x0 + bw, rh = y1 - y0 + bh;
    const size_t rB = align_up((size_t)rw * rh * 2, 256), cB = align_up((size_t)bw * bh * 2, 256);
    CallCtx *c = call_ctx(rB + cB + 512, rB + cB + 512);
    if (!c) return;
    pack2d(c->host, me->ref + (intptr_t)y0 * me->refStride + x0, me->refStride, rw, rh, 2);
    pack2d(c->host + rB, me->input, me->inputStride, bw, bh, 2);
    memcpy(c->host + rB + cB, cand, sizeof(cand));
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, rB + cB + sizeof(cand), hipMemcpyHostToDevice, c->stream))) return;
    uint32_t *dout = (uint32_t *)(c->dev + rB + cB + 256);
    k_dia_one<<<1, 256, 0, c->stream>>>((uint16_t *)c->dev, rw, x0, y0, (uint16_t *)(c->dev + rB), bw, bh,
                                        (int16_t *)(c->dev + rB + cB), (int)me->candNum, mnx, mny, mxx, mxy, cx, cy,
                                        me->searchWidth, me->searchHeight, me->sadLambda, me->predMv.mvX,
                                        me->predMv.mvY, dout);
    if (!XG_HIP(hipGetLastError())) return;
    uint32_t *hout = (uint32_t *)(c->host + rB + cB + 256);
    if (!XG_HIP(hipMemcpyAsync(hout, dout, 16, hipMemcpyDeviceToHost, c->stream))) return;
    if (!XG_HIP(hipStreamSynchronize(c->stream))) return;
    me->bestMv.mvX = (int16_t)(4 * (int)hout[0]);
    me->bestMv.mvY = (int16_t)(4 * (int)hout[1]);
    me->bestCost = hout[2];
    me->bestSad = hout[3];
}

void Xg266FullSubPelMe(xg_me_struct *me)
{
    if (!me || !me->input || !me->ref || me->width == 0 || me->height == 0 || me->width > 128 || me->height > 128) {
        set_error(XG_ERR_ARG, "Xg266FullSubPelMe: bad arguments");
        return;
    }
    const int qx = me->bestMv.mvX, qy = me->bestMv.mvY;
    // h266_sub_pel_refine.c:125-128: leave the MV alone when it is out of range
    if ((qx >> 2) < me->minMv.mvX || (qy >> 2) < me->minMv.mvY || (qx >> 2) > me->maxMv.mvX ||
        (qy >> 2) > me->maxMv.mvY)
        return;
    const int bw = (int)me->width, bh = (int)me->height;
    // reads span integer positions (qx>>2)-1 .. (qx>>2)+1 plus the 8-tap reach
    const int rx0 = (qx >> 2) - 1 - 3, ry0 = (qy >> 2) - 1 - 3;
    const int rw = bw + 10, rh = bh + 10;
    const size_t rB = align_up((size_t)rw * rh * 2, 256), cB = align_up((size_t)bw * bh * 2, 256);
    CallCtx *c = call_ctx(rB + cB + 256, rB + cB + 256);
    if (!c) return;
    pack2d(c->host, me->ref + (intptr_t)ry0 * me->refStride + rx0, me->refStride, rw, rh, 2);
    pack2d(c->host + rB, me->input, me->inputStride, bw, bh, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, rB + cB, hipMemcpyHostToDevice, c->stream))) return;
    uint32_t *dout = (uint32_t *)(c->dev + rB + cB);
    k_subpel_one<<<1, 256, 0, c->stream>>>((uint16_t *)c->dev, rw, rx0, ry0, (uint16_t *)(c->dev + rB), bw, bh,
                                           qx, qy, me->bestCost, me->predMv.mvX, me->predMv.mvY, me->sadLambda, dout);
    if (!XG_HIP(hipGetLastError())) return;
    uint32_t *hout = (uint32_t *)(c->host + rB + cB);
    if (!XG_HIP(hipMemcpyAsync(hout, dout, 12, hipMemcpyDeviceToHost, c->stream))) return;
    if (!XG_HIP(hipStreamSynchronize(c->stream))) return;
    me->bestMv.mvX = (int16_t)hout[0];
    me->bestMv.mvY = (int16_t)hout[1];
    me->bestCost = hout[2];
}

int32_t xg_search_blocks(xg_session *s, int32_t range, uint32_t sadLambda, int32_t ctuSize)
{
    SessionView v;
    if (!session_view(s, &v) || range < 0 || range > 32 || ctuSize < 4 || ctuSize > 256) {
        set_error(XG_ERR_ARG, "xg_search_blocks: bad arguments");
        return XG_ERR_ARG;
    }
    if (v.nblocks == 0) return XG_OK;
    int maxw = 0, maxh = 0;
    for (int i = 0; i < v.nblocks; i++) {
        const xg_block &B = v.hBlocks[i];
        if (B.pic >= kSessionPicSlots || B.ref >= kSessionPicSlots || !v.picSet[B.pic] || !v.picSet[B.ref]) {
            set_error(XG_ERR_ARG, "xg_search_blocks: block %d picture slot not set", i);
            return XG_ERR_ARG;
        }
        const xg_picture &P = v.hPics[B.pic], &R = v.hPics[B.ref];
        if (B.x < 0 || B.y < 0 || B.x + B.w > P.plane[0].width || B.y + B.h > P.plane[0].height ||
            R.plane[0].width != P.plane[0].width || R.plane[0].height != P.plane[0].height ||
            R.plane[0].pad < ctuSize + 16) {
            set_error(XG_ERR_ARG, "xg_search_blocks: block %d outside its picture or reference under-padded", i);
            return XG_ERR_ARG;
        }
        // the packed core reads the current block as dword pairs
        if ((B.w & (B.w - 1)) || B.w < 4 || B.w > 64 || B.h < 1 || B.h > 64 || (B.x & 1) || (P.plane[0].stride & 1) ||
            ((uintptr_t)P.plane[0].data & 3)) {
            set_error(XG_ERR_ARG, "xg_search_blocks: block %d: width must be 4..64 (a power of two) at an even x "
                      "in a dword-aligned plane", i);
            return XG_ERR_ARG;
        }
        maxw = std::max(maxw, (int)B.w);
        maxh = std::max(maxh, (int)B.h);
    }
    const int ws = packed_ws(2 * range + 1, maxw);
    const size_t lds = ((size_t)(2 * range + maxh) * ws + (size_t)maxw * maxh) * 2 + 16;
    if (lds > 64 * 1024) {
        set_error(XG_ERR_ARG, "xg_search_blocks: window exceeds the LDS budget");
        return XG_ERR_ARG;
    }
    int ev[2];
    const bool timed = session_timer_begin(s, 3, ev);
    k_search_blocks<<<v.nblocks, packed_threads(2 * range + 1), lds, v.stream>>>(v.dPics, v.dBlocks, range,
                                                                                sadLambda, ctuSize, ws);
    if (timed) session_timer_end(s, 3, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

int32_t xg_full_search_frame(xg_session *s, const xg_plane *cur, const xg_plane *ref, int32_t bw, int32_t bh,
                             int32_t range, uint32_t sadLambda, int32_t ctuSize, int16_t *bestMv,
                             uint32_t *bestCost, uint32_t *bestSad)
{
    if (!s) {
        set_error(XG_ERR_ARG, "xg_full_search_frame: bad arguments");
        return XG_ERR_ARG;
    }
    return xg::full_search_frame_on((hipStream_t)xg_session_stream(s), s, cur, ref, bw, bh, range, sadLambda,
                                    ctuSize, bestMv, bestCost, bestSad);
}

}  // extern "C"

int32_t xg::full_search_frame_on(hipStream_t st, xg_session *timer, const xg_plane *cur, const xg_plane *ref,
                                 int32_t bw, int32_t bh, int32_t range, uint32_t sadLambda, int32_t ctuSize,
                                 int16_t *bestMv, uint32_t *bestCost, uint32_t *bestSad)
{
    if (!cur || !ref || !bestMv || !bestCost || !bestSad || bw < 4 || bh < 4 || bw > 64 || bh > 64 ||
        range < 0 || range > 64 || cur->width != ref->width || cur->height != ref->height ||
        ref->pad < range + 4) {
        set_error(XG_ERR_ARG, "xg_full_search_frame: bad arguments");
        return XG_ERR_ARG;
    }
    const int nbx = cur->width / bw, nby = cur->height / bh;
    if (nbx <= 0 || nby <= 0) return XG_OK;
    const size_t lds = ((size_t)(2 * range + bw) * (2 * range + bh) + (size_t)bw * bh) * 2;
    if (lds > 64 * 1024) {
        set_error(XG_ERR_ARG, "xg_full_search_frame: window exceeds the LDS budget");
        return XG_ERR_ARG;
    }
    // packed form: widths 8, 16, 32 with a dword-aligned current plane
    const int nxm = 2 * range + 1, wsP = packed_ws(nxm, bw, kFrameK);
    const size_t ldsP = ((size_t)(2 * range + bh) * wsP + (size_t)bw * bh) * 2 + 16;
    const bool packed = (bw == 8 || bw == 16 || bw == 32) && ((uintptr_t)cur->data & 3) == 0 && (cur->stride & 1) == 0 &&
                        ldsP <= 64 * 1024;
    FsFrameArgs a;
    a.cur = cur->data; a.cs = cur->stride;
    a.ref = ref->data; a.rs = ref->stride;
    a.fw = cur->width; a.fh = cur->height; a.bw = bw; a.bh = bh; a.range = range; a.ctu = ctuSize; a.nbx = nbx;
    a.lambda = sadLambda;
    a.bestMv = bestMv; a.bestCost = bestCost; a.bestSad = bestSad;
    int ev[2];
    const bool timed = timer && session_timer_begin(timer, 3, ev);
    if (packed) {
        const int threads = packed_threads(nxm, kFrameK);
        if (bw == 8)
            k_full_search_frame_p<8><<<nbx * nby, threads, ldsP, st>>>(a, wsP);
        else if (bw == 16)
            k_full_search_frame_p<16><<<nbx * nby, threads, ldsP, st>>>(a, wsP);
        else
            k_full_search_frame_p<32><<<nbx * nby, threads, ldsP, st>>>(a, wsP);
    } else {
        k_full_search_frame<<<nbx * nby, 256, lds, st>>>(a);
    }
    if (timed) session_timer_end(timer, 3, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

// ---------------------------------------------------------------------------
// Serving search by CTU group
//
// A workgroup and a window staging per block would spend most of the time on
// setup at 1080p, where most blocks are 4..16 wide.  Here a workgroup of four
// waves takes a group (up to 64 blocks of one CTU): it takes every block's
// predictor (written by the AMVP stage, k_amvp_multi) and window, stages the
// union of the windows once in LDS (with the CTU's current samples), and
// spreads the blocks' items over its waves, 64 consecutive items of the
// group's list per wave step.  An item is K = 10 candidate columns of two
// candidate rows (blocks up to 16 wide: each window row read serves both) or
// of one row (wider blocks).  Each block's minimum (cost << 32 | raster index
// in its own window) is kept by an LDS 64-bit atomic min, so the winner and
// the tie order are those of search_block.  The LDS window starts at an even
// column; a block whose window starts at an odd one reads from one column to
// the left and skips its first candidate (s = 1).  A group whose union does
// not fit stages one block's window at a time instead.
// ---------------------------------------------------------------------------
namespace xg {

constexpr int kGroupWaves = 4;
constexpr int kGroupCurDw = 64 * 32;             // the group's current samples: 64 rows x 32 dwords
static_assert(kGroupCurDw % (64 * kGroupWaves) == 0, "whole current-sample rounds per thread");
constexpr int kGroupK = 10;                      // candidates per lane (an odd number of dwords apart: LDS banks)

struct GroupBlk {
    int b;                                       // block index in the list
    int x0, y0, nx, ny;                          // window in integer pel
    int ox, oy;                                  // window origin in the LDS window (ox even)
    int s;                                       // 1 when the origin is one column left of x0
    int nq, np, items, w, h;                     // column groups, row groups (of rpi rows), nq * np
    int rpi;                                     // candidate rows per item: 2 up to 16 wide, else 1
    int pmx, pmy;                                // predictor, quarter pel
    int cur;                                     // dword offset of the block in the current samples
};

// Dword offset of row R of an LDS window whose rows are ws2 dwords apart plus
// one dword after every second row: two rows apart is then an odd number of
// dwords, so lanes two candidate rows apart use different banks.
__device__ __forceinline__ int win_row(int R, int ws2) { return R * ws2 + (R >> 1); }

// K SADs (candidates K q .. K q + K - 1 of origin-relative column) of one
// candidate row for a block BW wide, h rows: window rows R .. R + h - 1 from
// dword column col (win_row layout, pitch ws2), current rows 32 dwords apart.
template <int BW>
__device__ __forceinline__ void group_item(const uint32_t *win, int col, int R, int ws2, const uint32_t *cp, int h,
                                           uint32_t acc[kGroupK])
{
    constexpr int ND = BW / 2, KD = kGroupK / 2;
#pragma unroll
    for (int k = 0; k < kGroupK; k++) acc[k] = 0;
    int off = win_row(R, ws2) + col;
#pragma unroll(BW >= 32 ? 1 : 2)
    for (int i = 0; i < h; i++) {
        uint32_t D[ND + KD], F[ND + KD - 1], C[ND];
#pragma unroll
        for (int j = 0; j < ND + KD; j++) D[j] = win[off + j];
#pragma unroll
        for (int j = 0; j < ND; j++) C[j] = cp[i * 32 + j];
#pragma unroll
        for (int j = 0; j < ND + KD - 1; j++) F[j] = align16(D[j + 1], D[j], 16);
#pragma unroll
        for (int j = 0; j < ND; j++)
#pragma unroll
            for (int m = 0; m < KD; m++) {
                acc[2 * m] = __builtin_amdgcn_sad_u16(D[j + m], C[j], acc[2 * m]);
                acc[2 * m + 1] = __builtin_amdgcn_sad_u16(F[j + m], C[j], acc[2 * m + 1]);
            }
        off += ws2 + ((R + i) & 1);
    }
}

// SADs of one window row against one current row, added to acc.
template <int ND>
__device__ __forceinline__ void row_sads(const uint32_t *D, const uint32_t *F, const uint32_t *C, uint32_t *acc)
{
    constexpr int KD = kGroupK / 2;
#pragma unroll
    for (int j = 0; j < ND; j++)
#pragma unroll
        for (int m = 0; m < KD; m++) {
            acc[2 * m] = __builtin_amdgcn_sad_u16(D[j + m], C[j], acc[2 * m]);
            acc[2 * m + 1] = __builtin_amdgcn_sad_u16(F[j + m], C[j], acc[2 * m + 1]);
        }
}

// The K SADs of two candidate rows at once (acc0: window rows R .., acc1: the
// next one): window row R + i serves current row i of the first and current
// row i - 1 of the second, so each window row is read and realigned once for
// both.  Reads window rows R .. R + h.
template <int BW>
__device__ __forceinline__ void group_item2(const uint32_t *win, int col, int R, int ws2, const uint32_t *cp, int h,
                                            uint32_t acc0[kGroupK], uint32_t acc1[kGroupK])
{
    constexpr int ND = BW / 2, KD = kGroupK / 2;
#pragma unroll
    for (int k = 0; k < kGroupK; k++) acc0[k] = acc1[k] = 0;
    uint32_t D[ND + KD], F[ND + KD - 1], C[ND], Cp[ND];
    int off = win_row(R, ws2) + col;
    auto load_row = [&](int i) {
#pragma unroll
        for (int j = 0; j < ND + KD; j++) D[j] = win[off + j];
#pragma unroll
        for (int j = 0; j < ND + KD - 1; j++) F[j] = align16(D[j + 1], D[j], 16);
        off += ws2 + ((R + i) & 1);
    };
    load_row(0);
#pragma unroll
    for (int j = 0; j < ND; j++) C[j] = cp[j];
    row_sads<ND>(D, F, C, acc0);
#pragma unroll 2
    for (int i = 1; i < h; i++) {
#pragma unroll
        for (int j = 0; j < ND; j++) Cp[j] = C[j];
        load_row(i);
#pragma unroll
        for (int j = 0; j < ND; j++) C[j] = cp[i * 32 + j];
        row_sads<ND>(D, F, C, acc0);
        row_sads<ND>(D, F, Cp, acc1);
    }
    load_row(h);
    row_sads<ND>(D, F, C, acc1);
}

// Runs the items of blocks [j0, j1), flattened: block j holds items
// ibeg[j] .. ibeg[j + 1] - 1 of the group, and each wave takes 64
// consecutive items at a time (stride kGroupWaves waves), so a wave may span
// the end of one block and the start of the next (the host orders a group's
// blocks by width, so such a wave mostly runs one width).
__device__ __forceinline__ void group_units(const uint32_t *win, int ws, const uint32_t *cur, const GroupBlk *gb,
                                            const int *ibeg, int j0, int j1, unsigned long long *gkey,
                                            uint32_t lambda)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int first = ibeg[j0], total = ibeg[j1] - first;
    for (int u0 = 64 * wave; u0 < total; u0 += 64 * kGroupWaves) {
        const int ig = first + u0 + lane;
        int lo = j0, hi = j1 - 1;                 // the block whose items hold ig
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ibeg[mid] <= ig) lo = mid;
            else hi = mid - 1;
        }
        const GroupBlk &G = gb[lo];
        const int it = ig - ibeg[lo];
        unsigned long long key = ~0ull;
        if (ig < first + total) {
            // item: candidate rows rpi p .. rpi p + rpi - 1, columns K q .. K q + K - 1
            // row group fastest: adjacent lanes read window rows rpi apart (win_row: distinct banks)
            const int q = it / G.np, p = it - q * G.np, dyi = G.rpi * p;
            const int R = G.oy + dyi, col = (G.ox >> 1) + (kGroupK / 2) * q;
            const uint32_t *cp = cur + G.cur;
            const int ws2 = ws >> 1;
            uint32_t acc[2][kGroupK];
            switch (G.w) {
            case 4: group_item2<4>(win, col, R, ws2, cp, G.h, acc[0], acc[1]); break;
            case 8: group_item2<8>(win, col, R, ws2, cp, G.h, acc[0], acc[1]); break;
            case 16: group_item2<16>(win, col, R, ws2, cp, G.h, acc[0], acc[1]); break;
            case 32: group_item<32>(win, col, R, ws2, cp, G.h, acc[0]); break;
            default: {                                   // 64 wide: two 32-wide halves
                uint32_t t[kGroupK];
                group_item<32>(win, col, R, ws2, cp, G.h, acc[0]);
                group_item<32>(win, col + 16, R, ws2, cp + 16, G.h, t);
#pragma unroll
                for (int k = 0; k < kGroupK; k++) acc[0][k] += t[k];
                break;
            }
            }
            // the lane's best: its candidates come in raster order (rows, then columns), so a
            // strict < keeps the lowest index among equal costs, as the 64-bit key order does
            uint32_t bc = 0xFFFFFFFFu, bi = 0;
            uint32_t bxl[kGroupK];
            bool vx[kGroupK];
#pragma unroll
            for (int k = 0; k < kGroupK; k++) {
                const int dx = kGroupK * q + k - G.s;
                vx[k] = dx >= 0 && dx < G.nx;
                bxl[k] = eg_bits(4 * (G.x0 + dx) - G.pmx);
            }
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int dy = dyi + r;
                if (r < G.rpi && dy < G.ny) {
                    const uint32_t by = eg_bits(4 * (G.y0 + dy) - G.pmy);
                    const int row0 = dy * G.nx + kGroupK * q - G.s;
#pragma unroll
                    for (int k = 0; k < kGroupK; k++) {
                        const uint32_t cost = (acc[r][k] << 8) + (bxl[k] + by) * lambda;
                        if (vx[k] && cost < bc) {
                            bc = cost;
                            bi = (uint32_t)(row0 + k);
                        }
                    }
                }
            }
            if (bc != 0xFFFFFFFFu) key = ((unsigned long long)bc << 32) | bi;
        }
        if (__builtin_amdgcn_readfirstlane(lo) == __builtin_amdgcn_readlane(lo, 63)) {
            // one block in the wave (blocks are ordered, so lanes 0 and 63 bound it)
#pragma unroll
            for (int m = 1; m < 64; m <<= 1) {
                const unsigned long long o = (unsigned long long)__shfl_xor((long long)key, m, 64);
                key = o < key ? o : key;
            }
            if (lane == 0) atomicMin(&gkey[lo], key);
        } else if (key != ~0ull) {
            atomicMin(&gkey[lo], key);
        }
    }
}

// rows x (ws/2) dwords of the reference luma plane from picture column cx
// (even) and row cy into win, in the win_row layout.  A thread's dwords i,
// i + blockDim, ... go in batches of 8 loads before their stores (a load per
// iteration would wait out each load's latency in turn); its (row, column)
// steps by (blockDim / nd, blockDim % nd) with a carry, with no division per
// dword.
__device__ __forceinline__ void stage_win_dw(uint32_t *win, int ws, const xg_plane &pl, int cx, int cy, int rows)
{
    const int nd = ws >> 1, n = rows * nd, nt = (int)blockDim.x;
    const int dr = nt / nd, dc = nt - dr * nd;
    const uint16_t *base = pl.data + (intptr_t)cy * pl.stride + cx;
    int r = (int)threadIdx.x / nd, c = (int)threadIdx.x - r * nd;
    for (int i0 = threadIdx.x; i0 < n; i0 += 8 * nt) {
        uint32_t v[8];
        int at[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool in = i0 + k * nt < n;
            at[k] = in ? win_row(r, nd) + c : -1;
            if (in) v[k] = *(const gdword *)(base + (intptr_t)r * pl.stride + 2 * c);
            r += dr;
            c += dc;
            if (c >= nd) {
                c -= nd;
                r++;
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (at[k] >= 0) win[at[k]] = v[k];
    }
}

__global__ void __launch_bounds__(64 * kGroupWaves) __attribute__((amdgpu_waves_per_eu(4))) k_search_groups(const xg_picture *pics,
                                                                    const SearchList *__restrict__ lists, int nLists,
                                                                    int range, uint32_t lambda, int ctu, int winDw)
{
    extern __shared__ uint32_t win[];                        // winDw dwords
    __shared__ uint32_t cur[kGroupCurDw];
    __shared__ GroupBlk gb[kSearchGroupMax];
    __shared__ unsigned long long gkey[kSearchGroupMax];
    __shared__ int ibeg[kSearchGroupMax + 1];
    __shared__ int box[6];                                   // union x0, y0, x1, y1; current x0, y0
    int l = 0;
    while (l + 1 < nLists && (int)blockIdx.x >= lists[l + 1].begin) l++;
    const SearchList &L = lists[l];
    const int2 G = L.groups[(int)blockIdx.x - L.begin];
    const int count = G.y, t = threadIdx.x;
    const xg_picture &P = pics[L.curSlot], &R = pics[L.refSlot];
    const int fw = P.plane[0].width, fh = P.plane[0].height;
    if (t == 0) {
        box[0] = box[1] = INT_MAX;
        box[2] = box[3] = INT_MIN;
        box[4] = box[5] = INT_MAX;
    }
    __syncthreads();
    xg_block B;
    int x0 = 0, y0 = 0, nx = 0, ny = 0;
    if (t < count) {
        const int b = L.gidx[G.x + t];
        B = L.blocks[b];
        B.pic = L.curSlot;
        B.ref = L.refSlot;
        const int mnx = (int16_t)(-B.x - ctu - 8), mny = (int16_t)(-B.y - ctu - 8);
        const int mxx = (int16_t)(fw + (ctu - B.w + 8) - B.x), mxy = (int16_t)(fh + (ctu - B.h + 8) - B.y);
        const int cx = min(max(((B.mvpx >> 2) + 2) >> 2, mnx), mxx), cy = min(max(((B.mvpy >> 2) + 2) >> 2, mny), mxy);
        x0 = max(cx - range, mnx);
        y0 = max(cy - range, mny);
        nx = min(cx + range, mxx) - x0 + 1;
        ny = min(cy + range, mxy) - y0 + 1;
        const int sx = (B.x + x0) & 1;                       // origin one column left when odd
        const int nq = (nx + sx + kGroupK - 1) / kGroupK;
        GroupBlk g;
        g.b = b;
        g.x0 = x0; g.y0 = y0; g.nx = nx; g.ny = ny;
        g.s = sx;
        g.nq = nq;
        g.rpi = B.w <= 16 ? 2 : 1;
        g.np = (ny + g.rpi - 1) / g.rpi;
        g.items = g.np * nq;
        g.w = B.w; g.h = B.h;
        g.pmx = B.mvpx >> 2; g.pmy = B.mvpy >> 2;
        gb[t] = g;
        gkey[t] = ~0ull;
        atomicMin(&box[0], B.x + x0 - sx);
        atomicMin(&box[1], B.y + y0);
        atomicMax(&box[2], B.x + x0 - sx + kGroupK * nq + B.w);   // columns the last item reads, exclusive
        atomicMax(&box[3], B.y + y0 + ny - 1 + B.h);
        atomicMin(&box[4], (int)B.x);
        atomicMin(&box[5], (int)B.y);
    }
    __syncthreads();
    // row pitch: an odd number of dwords, so the lanes' rows spread over the LDS banks
    const int ux0 = box[0], uy0 = box[1], rows = box[3] - box[1];
    const int ws = (((box[2] - box[0]) + 1) & ~1) | 2;
    const int cx0 = box[4], cy0 = box[5];
    if (t < count) {
        gb[t].ox = B.x + x0 - gb[t].s - ux0;
        gb[t].oy = B.y + y0 - uy0;
        gb[t].cur = (B.y - cy0) * 32 + ((B.x - cx0) >> 1);
    }
    if (t == 0) {
        ibeg[0] = 0;
        for (int j = 0; j < count; j++) ibeg[j + 1] = ibeg[j] + gb[j].items;
    }
    // the group's current samples (within 64x64 of (cx0, cy0): checked on the
    // host), every load of a thread issued before its first store
    {
        constexpr int kPer = kGroupCurDw / (64 * kGroupWaves);
        uint32_t v[kPer];
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int i = t + k * 64 * kGroupWaves, r = i >> 5, c = i & 31;
            const bool in = cy0 + r < fh && cx0 + 2 * c < fw;
            v[k] = in ? *(const gdword *)(P.plane[0].data + (intptr_t)(cy0 + r) * P.plane[0].stride + cx0 + 2 * c)
                      : 0u;
        }
#pragma unroll
        for (int k = 0; k < kPer; k++) cur[t + k * 64 * kGroupWaves] = v[k];
    }
    if (win_row(rows + 1, ws >> 1) <= winDw) {               // an item may read one row past the windows
        stage_win_dw(win, ws, R.plane[0], ux0, uy0, rows);
        __syncthreads();
        group_units(win, ws, cur, gb, ibeg, 0, count, gkey, lambda);
    } else {
        // one block's window at a time
        for (int j = 0; j < count; j++) {
            __syncthreads();                                 // the previous block's readers are done
            const GroupBlk g = gb[j];
            const int bx = ux0 + g.ox, byy = uy0 + g.oy;     // the block's window origin in the picture
            const int wsj = ((kGroupK * g.nq + g.w + 1) & ~1) | 2, rj = g.ny + g.h - 1;   // odd dword pitch
            stage_win_dw(win, wsj, R.plane[0], bx, byy, rj);
            if (t == 0) {
                gb[j].ox = 0;
                gb[j].oy = 0;
            }
            __syncthreads();
            group_units(win, wsj, cur, gb, ibeg, j, j + 1, gkey, lambda);
        }
    }
    __syncthreads();
    if (t < count) {
        const unsigned long long k = gkey[t];
        const int i = (int)(uint32_t)k;
        B.mvx = 16 * (x0 + i % nx);
        B.mvy = 16 * (y0 + i / nx);
        L.blocks[gb[t].b] = B;
    }
}

}  // namespace xg

void xg::search_groups(const xg_block *blocks, int n, std::vector<int> &gidx, std::vector<int2> &groups)
{
    // blocks of each 64x64 CTU (by top-left sample), in list order
    std::vector<std::pair<long, int>> key(n);
    for (int i = 0; i < n; i++) key[i] = {((long)(blocks[i].y >> 6) << 20) | (blocks[i].x >> 6), i};
    std::stable_sort(key.begin(), key.end(), [](const std::pair<long, int> &a, const std::pair<long, int> &b) {
        return a.first < b.first;
    });
    gidx.clear();
    groups.clear();
    for (int a = 0; a < n;) {
        int e = a;
        while (e < n && key[e].first == key[a].first) e++;
        int x0 = INT_MAX, y0 = INT_MAX, x1 = INT_MIN, y1 = INT_MIN;
        for (int k = a; k < e; k++) {
            const xg_block &B = blocks[key[k].second];
            x0 = std::min(x0, (int)B.x);
            y0 = std::min(y0, (int)B.y);
            x1 = std::max(x1, B.x + B.w);
            y1 = std::max(y1, B.y + B.h);
        }
        const bool fits = x1 - x0 <= 64 && y1 - y0 <= 64;
        // narrowest first: the waves of k_search_groups run consecutive items
        std::stable_sort(key.begin() + a, key.begin() + e, [&](const std::pair<long, int> &p, const std::pair<long, int> &q) {
            return blocks[p.second].w < blocks[q.second].w;
        });
        for (int k = a; k < e;) {
            const int c = fits ? std::min(kSearchGroupMax, e - k) : 1;   // blocks past 64x64: one per group
            groups.push_back(make_int2((int)gidx.size(), c));
            for (int m = 0; m < c; m++) gidx.push_back(key[k + m].second);
            k += c;
        }
        a = e;
    }
}

int32_t xg::search_groups_enqueue(hipStream_t st, xg_session *timer, const xg_picture *dPics, const SearchList *dLists,
                                  int nLists, int nGroups, int range, uint32_t lambda, int ctu)
{
    if (nLists <= 0 || nGroups <= 0) return XG_OK;
    // room for one 64x64 block's window at this range (the fallback), and for
    // the union of a CTU's windows when their predictors agree
    const int nqMax = (2 * range + 2 + kGroupK - 1) / kGroupK;
    // row pitches plus one dword per two rows (win_row)
    const int oneRows = 2 * range + 65, oneNd = ((kGroupK * nqMax + 64) + 1) / 2 + 1;
    const int one = oneRows * oneNd + oneRows / 2;
    const int uRows = 2 * range + 64 + 4, uNd = (2 * range + 64 + kGroupK + 8) / 2 + 1;
    const int winDw = std::max(one, uRows * uNd + uRows / 2);
    if (winDw * 4 + (int)(kGroupCurDw * 4 + kSearchGroupMax * (sizeof(GroupBlk) + 12)) > 64 * 1024) {
        set_error(XG_ERR_ARG, "search_groups_enqueue: window exceeds the LDS budget");
        return XG_ERR_ARG;
    }
    int ev[2];
    const bool timed = timer && session_timer_begin(timer, 3, ev);
    k_search_groups<<<nGroups, 64 * kGroupWaves, (size_t)winDw * 4, st>>>(dPics, dLists, nLists, range, lambda, ctu,
                                                                          winDw);
    if (timed) session_timer_end(timer, 3, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

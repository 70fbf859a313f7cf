// xg_dmvr.hip -- decoder-side MV refinement (DMVR) of bi-predicted blocks.
//
// Per block (both lists' MVs, 1/16 pel) the refinement of
// h266_encoder/source/h266_inter_prediction.c:976-1195:
//   * each list's MV is clipped to the CU's integer-pel window, moved by
//     -2 pels in x and y, and the block plus 4 samples of padding is predicted
//     with the 2-tap bilinear filter of h266_core/source/h266_bilinear_filter.c
//     (10-bit: Copy keeps the samples, each filter pass is (a*(16-f) + b*f + 8) >> 4);
//   * per subblock of min(16, side), 25 mirrored offsets (+d on list 0, -d on
//     list 1) are costed with a SAD over every second row (the S16 SAD with
//     subShift 1, >> 1); the centre cost less a quarter both starts the scan
//     and, below width x height, ends the search early;
//   * the best offset, if inside the 3x3 interior, gets the parametric
//     sub-pel step (Xin266SubPelErrorSrfc, Xin266DivForMaxQ7).
// One 64-thread workgroup per subblock (each XCD's workgroups take one
// contiguous eighth of the subblocks): both lists' source regions (at most
// 21 x 21 samples) and bilinear predictions (20 x 20) live in LDS, the 25
// costs are summed row by row across the lanes and chosen by a wave minimum
// over packed keys; lane 0 runs the sub-pel step.
//
// The per-call table slots (bilinear filters, S16 SAD) and the per-call
// refinement stage host data and run the same device code.
#include <algorithm>
#include <climits>
#include <mutex>
#include <string.h>
#include <vector>

#include "xg_batch.hpp"
#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

struct DmvrPic {
    const uint16_t *data;   // sample (0, 0)
    intptr_t stride;
};

// One block as the plan keeps it on the device.
struct DmvrTask {
    int32_t x, y, lgW, lgH;
    int32_t mv[2][2];           // [list][x, y], 1/16 pel
    int32_t minX, minY, maxX, maxY;
    int32_t out;                // index of the block's first subblock
};

__constant__ int8_t c_dmvrSearch[25][2] = {
    {-2, -2}, {-1, -2}, {0, -2}, {1, -2}, {2, -2}, {-2, -1}, {-1, -1}, {0, -1}, {1, -1}, {2, -1},
    {-2, 0},  {-1, 0},  {0, 0},  {1, 0},  {2, 0},  {-2, 1},  {-1, 1},  {0, 1},  {1, 1},  {2, 1},
    {-2, 2},  {-1, 2},  {0, 2},  {1, 2},  {2, 2}};

__device__ __forceinline__ int bili2(int a, int b, int f) { return ((16 - f) * a + f * b + 8) >> 4; }

// Xin266DivForMaxQ7 (h266_inter_prediction.c:864-912)
__device__ int div_q7(long long N, long long D)
{
    const bool neg = N < 0;
    if (neg) N = -N;
    int q = 0;
    D <<= 3;
    if (N >= D) { N -= D; q++; }
    q <<= 1;
    D >>= 1;
    if (N >= D) { N -= D; q++; }
    q <<= 1;
    if (N >= (D >> 1)) q++;
    return neg ? -q : q;
}

// Xin266SubPelRefineDmvr / Xin266SubPelErrorSrfc (:914-974): c = centre, left, top, right, bottom
__device__ void dmvr_subpel(int &dx, int &dy, const long long c[5])
{
    if (abs(dx) == 32 || abs(dy) == 32) return;
#pragma unroll
    for (int hv = 0; hv < 2; hv++) {
        const long long num = (c[hv + 1] - c[hv + 3]) * 16;
        const long long den = c[hv + 1] + c[hv + 3] - 2 * c[0];
        if (den == 0) continue;
        const int d = (c[hv + 1] != c[0] && c[hv + 3] != c[0]) ? div_q7(num, den) : (c[hv + 1] == c[0] ? -8 : 8);
        if (hv == 0) dx = (int16_t)(dx + d);
        else dy = (int16_t)(dy + d);
    }
}

// integer origin and phase of one list's bilinear prediction (:1040-1056)
__device__ __forceinline__ void dmvr_origin(const DmvrTask &t, int l, int &ix, int &iy, int &fx, int &fy)
{
    const int mvx = min(max(t.mv[l][0], t.minX * 16), t.maxX * 16) - 32;
    const int mvy = min(max(t.mv[l][1], t.minY * 16), t.maxY * 16) - 32;
    const int ox = t.x * 16 + mvx, oy = t.y * 16 + mvy;
    fx = ox & 15;
    fy = oy & 15;
    ix = ox >> 4;
    iy = oy >> 4;
}

__global__ void __launch_bounds__(64) k_dmvr(DmvrPic p0, DmvrPic p1, const DmvrTask *__restrict__ tasks,
                                              const int2 *__restrict__ subs, int nSub, int16_t *__restrict__ mvd)
{
    __shared__ uint16_t src[2][21 * 21];
    __shared__ __align__(16) int16_t bil[2][20 * 20];
    __shared__ uint32_t cost[25];
    const int lane = threadIdx.x;
    // XCD-aware split: workgroup g runs on XCD g % 8, so each XCD takes one
    // contiguous eighth of the subblocks (neighbouring blocks, the same
    // reference rows) and its L2 fetches those rows once
    int first = blockIdx.x, end = nSub, step = gridDim.x;
    if (gridDim.x >= 8 && gridDim.x % 8 == 0) {
        const int xcd = (int)blockIdx.x % 8, per = (int)gridDim.x / 8;
        first = (int)((long long)nSub * xcd / 8) + (int)blockIdx.x / 8;
        end = (int)((long long)nSub * (xcd + 1) / 8);
        step = per;
    }
    for (int si = first; si < end; si += step) {
        const int2 m = subs[si];
        const DmvrTask t = tasks[m.x];
        const int W = 1 << t.lgW, H = 1 << t.lgH;
        const int sw = min(W, 16), sh = min(H, 16);
        const int perRow = W / sw;
        const int sx = (m.y % perRow) * sw, sy = (m.y / perRow) * sh;
        const int rw = sw + 5, rh = sh + 5;          // source region of the subblock (filter reach included)
        const int bw = sw + 4, bh = sh + 4;          // bilinear prediction of the subblock plus padding
        int fxy[2][2];
        // Both lists' source regions.  Lane = column (rw <= 21 < 32) + 32 *
        // row parity, so a lane holds at most kRowsPerLane rows of each list;
        // it issues all their loads before its first LDS store, and the
        // subblock waits for one memory latency instead of one per row.  A
        // walking pointer (hidden from the compiler by an empty asm) keeps
        // the row offsets out of scalar registers.
        constexpr int kRowsPerLane = 11;                 // ceil(21 / 2)
        const int c = lane & 31, r0 = lane >> 5;
        uint32_t v[2][kRowsPerLane];
#pragma unroll
        for (int l = 0; l < 2; l++) {
            int ix, iy, fx, fy;
            dmvr_origin(t, l, ix, iy, fx, fy);
            fxy[l][0] = fx;
            fxy[l][1] = fy;
            const DmvrPic p = l ? p1 : p0;
            const gword *g = (const gword *)(p.data + (intptr_t)(iy + sy + r0) * p.stride + (ix + sx + c));
            const intptr_t step = 2 * p.stride;
            // the last row and column are read only by the filters that need them
            const bool col = c < rw && (c < bw || fx);
#pragma unroll
            for (int j = 0; j < kRowsPerLane; j++) {
                const int r = r0 + 2 * j;
                v[l][j] = 0;
                if (col && r < rh && (r < bh || fy)) v[l][j] = *g;
                g += step;
                asm volatile("" : "+v"(g));
            }
        }
        if (lane < 25) cost[lane] = 0;
#pragma unroll
        for (int l = 0; l < 2; l++)
#pragma unroll
            for (int j = 0; j < kRowsPerLane; j++) {
                const int r = r0 + 2 * j;
                if (c < rw && r < rh) src[l][r * 21 + c] = (uint16_t)v[l][j];
            }
        __syncthreads();
        // bilinear predictions, all of a lane's rows unrolled (bh <= 20)
#pragma unroll
        for (int l = 0; l < 2; l++) {
            const int fx = fxy[l][0], fy = fxy[l][1];
            if (c < bw)                                   // bw <= 20 < 32
#pragma unroll
                for (int j = 0; j < kRowsPerLane - 1; j++) {
                    const int r = r0 + 2 * j;
                    if (r >= bh) break;
                    const uint16_t *q = &src[l][r * 21 + c];
                    int b;
                    if (fx && fy) b = bili2(bili2(q[0], q[1], fx), bili2(q[21], q[22], fx), fy);
                    else if (fx) b = bili2(q[0], q[1], fx);
                    else if (fy) b = bili2(q[0], q[21], fy);
                    else b = q[0];
                    bil[l][r * 20 + c] = (int16_t)b;
                }
        }
        __syncthreads();
        // 25 offsets x sh/2 rows: each item is one row's SAD on packed sample
        // pairs (v_sad_u16; both offsets' rows start at the same parity, so one
        // realignment shift serves both), summed into its offset's cost
        {
            const uint32_t *B0 = reinterpret_cast<const uint32_t *>(bil[0]);
            const uint32_t *B1 = reinterpret_cast<const uint32_t *>(bil[1]);
            const int nr = sh >> 1, lnr = __builtin_ctz(nr);  // sh is a power of two
            for (int it = lane; it < 25 * nr; it += 64) {
                const int i = it >> lnr, r = 2 * (it & (nr - 1));
                const int dx = c_dmvrSearch[i][0], dy = c_dmvrSearch[i][1];
                const int ia = (2 + dy + r) * 20 + 2 + dx, ib = (2 - dy + r) * 20 + 2 - dx;
                const uint32_t sh16 = (ia & 1) * 16;
                const uint32_t *pa = B0 + (ia >> 1), *pb = B1 + (ib >> 1);
                uint32_t acc = 0;
                for (int k = 0; k < (sw >> 1); k++)
                    acc = __builtin_amdgcn_sad_u16(align16(pa[k + 1], pa[k], sh16), align16(pb[k + 1], pb[k], sh16),
                                                   acc);
                atomicAdd(&cost[i], acc);
            }
        }
        __syncthreads();
        // the scan of h266_inter_prediction.c:1125-1140 as one wave minimum:
        // key (cost << 5 | code), the centre (cost less a quarter) code 0 so
        // it keeps ties, offset i code i + 1 so the earliest of equal costs
        // wins (costs < 2^17: 8 rows of 16 differences of 10-bit samples)
        const uint32_t adj = cost[12] - (cost[12] >> 2);
        uint32_t key = 0xFFFFFFFFu;
        if (lane < 25) key = lane == 12 ? adj << 5 : (cost[lane] << 5) | (uint32_t)(lane + 1);
#pragma unroll
        for (int m = 1; m < 32; m <<= 1) key = min(key, (uint32_t)__shfl_xor((int)key, m, 64));
        if (lane == 0) {
            int dx = 0, dy = 0;
            cost[12] = adj;
            if (adj >= (uint32_t)(sw * sh)) {
                const int best = (key & 31) ? (int)(key & 31) - 1 : 12;
                dx = 16 * c_dmvrSearch[best][0];
                dy = 16 * c_dmvrSearch[best][1];
                if (abs(dx) < 32 && abs(dy) < 32) {
                    const long long c[5] = {cost[best], cost[best - 1], cost[best - 5], cost[best + 1], cost[best + 5]};
                    dmvr_subpel(dx, dy, c);
                }
            }
            mvd[2 * (t.out + m.y)] = (int16_t)dx;
            mvd[2 * (t.out + m.y) + 1] = (int16_t)dy;
        }
        __syncthreads();
    }
}

// Per-call bilinear filters on a packed staged region: kind 0 Copy, 1 Hor,
// 2 Vet, 3 HorVet; the source is (h + 1) x (w + 5) packed at stride w + 5.
__global__ void k_bili(const uint16_t *__restrict__ s, int kind, int fx, int fy, int w, int h,
                       int16_t *__restrict__ d)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c >= w + 4 || r >= h) return;
    const int ss = w + 5;
    const uint16_t *q = s + r * ss + c;
    int v;
    if (kind == 3) v = bili2(bili2(q[0], q[1], fx), bili2(q[ss], q[ss + 1], fx), fy);
    else if (kind == 1) v = bili2(q[0], q[1], fx);
    else if (kind == 2) v = bili2(q[0], q[ss], fy);
    else v = q[0];
    d[r * (w + 4) + c] = (int16_t)v;
}

// Per-call S16 SAD over every 2^subShift-th row, scaled back by << subShift.
__global__ void k_sad_s16(const int16_t *__restrict__ a, const int16_t *__restrict__ b, int w, int h, int subShift,
                          uint32_t *__restrict__ out)
{
    uint32_t s = 0;
    const int rows = (h + (1 << subShift) - 1) >> subShift;
    for (int e = threadIdx.x; e < rows * w; e += 256) {
        const int r = (e / w) << subShift, c = e % w;
        s += (uint32_t)abs((int)a[r * w + c] - (int)b[r * w + c]);
    }
    s = block_sum256(s);
    if (threadIdx.x == 0) out[0] = s << subShift;
}

}  // namespace xg

using namespace xg;

namespace {
struct DmvrPlan {
    // per list, the extent of every block's source region (clipped MV, -2 pels,
    // (w + 5) x (h + 5)) and a block reaching each side: the run checks the
    // planes against it
    int lo[2][2], hi[2][2], loAt[2][2], hiAt[2][2];
    DmvrTask *dTasks = nullptr;
    int2 *dSubs = nullptr;
    int n = 0, nSub = 0, taskCap = 0, subCap = 0;
    bool valid = false;
};
std::mutex g_dplan_mu;
std::vector<std::pair<xg_session *, DmvrPlan>> g_dplans;

DmvrPlan *dplan_for(xg_session *s, bool create)
{
    for (auto &p : g_dplans)
        if (p.first == s) return &p.second;
    if (!create) return nullptr;
    g_dplans.push_back({s, DmvrPlan()});
    return &g_dplans.back().second;
}

int32_t dplan_fail(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_dplan_mu);
    if (DmvrPlan *p = dplan_for(s, false)) p->valid = false;
    return XG_ERR_ARG;
}

int lg_side(int v)
{
    if (v < 8 || v > 128 || (v & (v - 1))) return -1;
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

bool build_task(const xg_dmvr_block &b, int out, DmvrTask &t)
{
    t.lgW = lg_side(b.width);
    t.lgH = lg_side(b.height);
    if (t.lgW < 0 || t.lgH < 0 || b.minMvX > b.maxMvX || b.minMvY > b.maxMvY) return false;
    t.x = b.x;
    t.y = b.y;
    t.mv[0][0] = b.mv0x;
    t.mv[0][1] = b.mv0y;
    t.mv[1][0] = b.mv1x;
    t.mv[1][1] = b.mv1y;
    t.minX = b.minMvX;
    t.minY = b.minMvY;
    t.maxX = b.maxMvX;
    t.maxY = b.maxMvY;
    t.out = out;
    return true;
}

int subblocks(const DmvrTask &t)
{
    const int W = 1 << t.lgW, H = 1 << t.lgH;
    return (W / std::min(W, 16)) * (H / std::min(H, 16));
}
}  // namespace

extern "C" {

int32_t xg_dmvr_plan(xg_session *s, const xg_dmvr_block *blocks, int32_t n, int32_t *subOffset)
{
    if (!s || n < 0 || (n > 0 && !blocks)) {
        set_error(XG_ERR_ARG, "xg_dmvr_plan: bad arguments");
        return s ? dplan_fail(s) : XG_ERR_ARG;
    }
    std::vector<DmvrTask> tasks((size_t)n);
    std::vector<int2> subs;
    int out = 0;
    for (int i = 0; i < n; i++) {
        if (!build_task(blocks[i], out, tasks[i])) {
            set_error(XG_ERR_ARG, "xg_dmvr_plan: block %d: sides must be 8..128 and powers of two, window ordered", i);
            return dplan_fail(s);
        }
        if (subOffset) subOffset[i] = out;
        const int k = subblocks(tasks[i]);
        for (int j = 0; j < k; j++) subs.push_back(make_int2(i, j));
        out += k;
    }
    if (subOffset) subOffset[n] = out;
    std::lock_guard<std::mutex> g(g_dplan_mu);
    DmvrPlan *p = dplan_for(s, true);
    p->valid = false;
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    if (!grow(p->dTasks, p->taskCap, std::max(n, 1), st) || !grow(p->dSubs, p->subCap, std::max(out, 1), st))
        return XG_ERR_NOMEM;
    if (n && !XG_HIP(hipMemcpyAsync(p->dTasks, tasks.data(), sizeof(DmvrTask) * (size_t)n, hipMemcpyHostToDevice, st)))
        return XG_ERR_HIP;
    if (out && !XG_HIP(hipMemcpyAsync(p->dSubs, subs.data(), sizeof(int2) * (size_t)out, hipMemcpyHostToDevice, st)))
        return XG_ERR_HIP;
    if (!XG_HIP(hipStreamSynchronize(st))) return XG_ERR_HIP;
    for (int l = 0; l < 2; l++)
        for (int d = 0; d < 2; d++) {
            p->lo[l][d] = INT_MAX;
            p->hi[l][d] = INT_MIN;
            p->loAt[l][d] = p->hiAt[l][d] = 0;
        }
    for (int i = 0; i < n; i++) {
        const DmvrTask &t = tasks[(size_t)i];
        for (int l = 0; l < 2; l++) {
            const int mvx = std::min(std::max(t.mv[l][0], t.minX * 16), t.maxX * 16) - 32;
            const int mvy = std::min(std::max(t.mv[l][1], t.minY * 16), t.maxY * 16) - 32;
            const int o[2] = {(t.x * 16 + mvx) >> 4, (t.y * 16 + mvy) >> 4};
            const int e[2] = {o[0] + (1 << t.lgW) + 5, o[1] + (1 << t.lgH) + 5};
            for (int d = 0; d < 2; d++) {
                if (o[d] < p->lo[l][d]) p->lo[l][d] = o[d], p->loAt[l][d] = i;
                if (e[d] > p->hi[l][d]) p->hi[l][d] = e[d], p->hiAt[l][d] = i;
            }
        }
    }
    p->n = n;
    p->nSub = out;
    p->valid = true;
    return XG_OK;
}

// every list's source region inside the padded plane, from the plan's extents
static bool reads_inside(const DmvrPlan &p, const xg_plane *ref0, const xg_plane *ref1, int *bad)
{
    for (int l = 0; l < 2; l++) {
        const xg_plane *pl = l ? ref1 : ref0;
        const int size[2] = {pl->width, pl->height};
        for (int d = 0; d < 2; d++) {
            if (p.lo[l][d] < -pl->pad) {
                *bad = p.loAt[l][d];
                return false;
            }
            if (p.hi[l][d] > size[d] + pl->pad) {
                *bad = p.hiAt[l][d];
                return false;
            }
        }
    }
    return true;
}

int32_t xg_dmvr_run(xg_session *s, const xg_plane *ref0, const xg_plane *ref1, int16_t *mvd)
{
    if (!s || !ref0 || !ref1 || !ref0->data || !ref1->data || !mvd) {
        set_error(XG_ERR_ARG, "xg_dmvr_run: bad arguments");
        return XG_ERR_ARG;
    }
    DmvrPlan *p;
    {
        std::lock_guard<std::mutex> g(g_dplan_mu);
        p = dplan_for(s, false);
    }
    if (!p || !p->valid) {
        set_error(XG_ERR_STATE, "xg_dmvr_run: no plan (call xg_dmvr_plan first; a failed plan leaves none)");
        return XG_ERR_STATE;
    }
    if (p->nSub == 0) return XG_OK;
    int bad = -1;
    if (!reads_inside(*p, ref0, ref1, &bad)) {
        set_error(XG_ERR_ARG, "xg_dmvr_run: block %d reads outside the padded reference planes", bad);
        return XG_ERR_ARG;
    }
    static const int slots = resident_slots((const void *)k_dmvr, 64);
    DmvrPic a{ref0->data, ref0->stride}, b{ref1->data, ref1->stride};
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    int ev[2];
    const bool timed = session_timer_begin(s, 6, ev);
    int grid = std::min(p->nSub, slots);
    if (grid >= 8) grid -= grid % 8;                // whole XCD rounds (k_dmvr's split)
    k_dmvr<<<grid, 64, 0, st>>>(a, b, p->dTasks, p->dSubs, p->nSub, mvd);
    if (timed) session_timer_end(s, 6, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

// ---- per-call table slots ----
static void bili_call(int kind, const xg_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac, uint32_t w,
                      uint32_t h, const char *fn)
{
    if (w == 0 || h == 0 || w > 128 || h > 132) {
        set_error(XG_ERR_ARG, "%s: unsupported %ux%u", fn, w, h);
        return;
    }
    // exactly the samples the reference function reads: w + 4 (+1 with a
    // horizontal pass) columns, h (+1 with a vertical pass) rows
    const int cols = (int)w + 4 + (kind & 1), rows = (int)h + (kind >> 1);
    const size_t inB = align_up((size_t)(h + 1) * (w + 5) * 2, 256), outB = (size_t)h * (w + 4) * 2;
    CallCtx *c = call_ctx(inB + outB, inB + outB);
    if (!c) return;
    memset(c->host, 0, (size_t)(h + 1) * (w + 5) * 2);
    for (int r = 0; r < rows; r++) memcpy(c->host + (size_t)r * (w + 5) * 2, src + (intptr_t)r * ss, (size_t)cols * 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, (size_t)(h + 1) * (w + 5) * 2, hipMemcpyHostToDevice, c->stream)))
        return;
    dim3 grid((w + 4 + 63) / 64, h);
    k_bili<<<grid, 64, 0, c->stream>>>((const uint16_t *)c->dev, kind, frac & 15, frac >> 4, (int)w, (int)h,
                                       (int16_t *)(c->dev + inB));
    if (!XG_HIP(hipGetLastError())) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + inB, c->dev + inB, outB, hipMemcpyDeviceToHost, c->stream))) return;
    if (!XG_HIP(hipStreamSynchronize(c->stream))) return;
    unpack2d(dst, ds, c->host + inB, (int)w + 4, (int)h, 2);
}

void Xg266BiliInterpCopy(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride, int32_t frac,
                         uint32_t width, uint32_t height)
{
    bili_call(0, src, srcStride, dst, dstStride, frac, width, height, "Xg266BiliInterpCopy");
}

void Xg266BiliInterpHor(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride, int32_t frac,
                        uint32_t width, uint32_t height)
{
    bili_call(1, src, srcStride, dst, dstStride, frac, width, height, "Xg266BiliInterpHor");
}

void Xg266BiliInterpVet(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride, int32_t frac,
                        uint32_t width, uint32_t height)
{
    bili_call(2, src, srcStride, dst, dstStride, frac, width, height, "Xg266BiliInterpVet");
}

void Xg266BiliInterpHorVet(const xg_pixel *src, intptr_t srcStride, int16_t *dst, intptr_t dstStride, int32_t frac,
                           uint32_t width, uint32_t height)
{
    bili_call(3, src, srcStride, dst, dstStride, frac, width, height, "Xg266BiliInterpHorVet");
}

void Xg26xComputeSadS16(int16_t *input, intptr_t inputStride, int16_t *pred, intptr_t predStride, uint32_t width,
                        uint32_t height, uint32_t subShift, uint32_t *sad)
{
    *sad = UINT32_MAX;
    if (width == 0 || height == 0 || width > 128 || height > 128 || subShift > 4) {
        set_error(XG_ERR_ARG, "Xg26xComputeSadS16: unsupported %ux%u subShift %u", width, height, subShift);
        return;
    }
    const size_t n = (size_t)width * height, blk = align_up(n * 2, 256);
    CallCtx *c = call_ctx(2 * blk + 256, 2 * blk + 256);
    if (!c) return;
    pack2d(c->host, input, inputStride, (int)width, (int)height, 2);
    pack2d(c->host + blk, pred, predStride, (int)width, (int)height, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, 2 * blk, hipMemcpyHostToDevice, c->stream))) return;
    k_sad_s16<<<1, 256, 0, c->stream>>>((const int16_t *)c->dev, (const int16_t *)(c->dev + blk), (int)width,
                                        (int)height, (int)subShift, (uint32_t *)(c->dev + 2 * blk));
    if (!XG_HIP(hipGetLastError())) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + 2 * blk, c->dev + 2 * blk, 4, hipMemcpyDeviceToHost, c->stream))) return;
    if (!XG_HIP(hipStreamSynchronize(c->stream))) return;
    *sad = *(uint32_t *)(c->host + 2 * blk);
}

int32_t Xg266DeriveMvFromDmvr(const xg_dmvr *d, int16_t *mvdL0SubPu)
{
    if (!d || !d->ref0 || !d->ref1 || !mvdL0SubPu) {
        set_error(XG_ERR_ARG, "Xg266DeriveMvFromDmvr: bad arguments");
        return XG_ERR_ARG;
    }
    xg_dmvr_block b{d->cuPelX, d->cuPelY, (int32_t)d->width, (int32_t)d->height, d->mv0x, d->mv0y, d->mv1x, d->mv1y,
                    d->minMvX, d->minMvY, d->maxMvX, d->maxMvY};
    DmvrTask t;
    if (!build_task(b, 0, t)) {
        set_error(XG_ERR_ARG, "Xg266DeriveMvFromDmvr: unsupported %ux%u or window", d->width, d->height);
        return XG_ERR_ARG;
    }
    // each list's (h + 5) x (w + 5) source region, staged; the task is
    // rewritten so its origin is (0, 0) of the staged region, same phases
    const int W = (int)d->width, H = (int)d->height, rw = W + 5, rh = H + 5;
    const int nSub = subblocks(t);
    const size_t reg = align_up((size_t)rw * rh * 2, 256);
    const size_t bytes = 2 * reg + align_up(sizeof(DmvrTask), 256) + align_up(sizeof(int2) * nSub, 256) +
                         (size_t)nSub * 4;
    CallCtx *c = call_ctx(bytes, bytes);
    if (!c) return XG_ERR_HIP;
    const xg_pixel *refs[2] = {d->ref0, d->ref1};
    const intptr_t strides[2] = {d->stride0, d->stride1};
    DmvrTask local = t;
    local.x = local.y = 0;
    local.minX = local.minY = -16384;
    local.maxX = local.maxY = 16383;
    for (int l = 0; l < 2; l++) {
        const int mvx = std::min(std::max(t.mv[l][0], t.minX * 16), t.maxX * 16) - 32;
        const int mvy = std::min(std::max(t.mv[l][1], t.minY * 16), t.maxY * 16) - 32;
        const int ox = t.x * 16 + mvx, oy = t.y * 16 + mvy;
        const int fx = ox & 15, fy = oy & 15, ix = ox >> 4, iy = oy >> 4;
        // read what the reference reads: the last column / row only for a horizontal / vertical phase
        const int cols = W + 4 + (fx ? 1 : 0), rows = H + 4 + (fy ? 1 : 0);
        memset(c->host + l * reg, 0, (size_t)rw * rh * 2);
        for (int r = 0; r < rows; r++)
            memcpy(c->host + l * reg + (size_t)r * rw * 2, refs[l] + (intptr_t)(iy + r) * strides[l] + ix,
                   (size_t)cols * 2);
        local.mv[l][0] = fx + 32;
        local.mv[l][1] = fy + 32;
    }
    char *hp = c->host + 2 * reg;
    memcpy(hp, &local, sizeof(local));
    int2 *hs = (int2 *)(hp + align_up(sizeof(DmvrTask), 256));
    for (int j = 0; j < nSub; j++) hs[j] = make_int2(0, j);
    const size_t upB = 2 * reg + align_up(sizeof(DmvrTask), 256) + sizeof(int2) * nSub;
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, upB, hipMemcpyHostToDevice, c->stream))) return XG_ERR_HIP;
    const size_t outOff = 2 * reg + align_up(sizeof(DmvrTask), 256) + align_up(sizeof(int2) * nSub, 256);
    DmvrPic a{(const uint16_t *)c->dev, rw}, bb{(const uint16_t *)(c->dev + reg), rw};
    k_dmvr<<<nSub, 64, 0, c->stream>>>(a, bb, (const DmvrTask *)(c->dev + 2 * reg),
                                       (const int2 *)(c->dev + 2 * reg + align_up(sizeof(DmvrTask), 256)), nSub,
                                       (int16_t *)(c->dev + outOff));
    if (!XG_HIP(hipGetLastError())) return XG_ERR_HIP;
    if (!XG_HIP(hipMemcpyAsync(c->host + outOff, c->dev + outOff, (size_t)nSub * 4, hipMemcpyDeviceToHost, c->stream)))
        return XG_ERR_HIP;
    if (!XG_HIP(hipStreamSynchronize(c->stream))) return XG_ERR_HIP;
    memcpy(mvdL0SubPu, c->host + outOff, (size_t)nSub * 4);
    return XG_OK;
}

}  // extern "C"

void xg::dmvr_release(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_dplan_mu);
    for (size_t i = 0; i < g_dplans.size(); i++)
        if (g_dplans[i].first == s) {
            if (g_dplans[i].second.dTasks) (void)hipFree(g_dplans[i].second.dTasks);
            if (g_dplans[i].second.dSubs) (void)hipFree(g_dplans[i].second.dSubs);
            g_dplans.erase(g_dplans.begin() + (long)i);
            return;
        }
}

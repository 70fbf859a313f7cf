// xg_bipred.hip -- bi-prediction distortion of a frame's blocks.
//
// The B-frame step of h266_encoder/source/h266_analyze_inter_cu.c:1212-1222:
// after both lists' uni-prediction searches, the distortion of the input
// against the rounded average of the two predictions,
// pfXinComputeAvgDist(input, pred0, pred1) -- SATD or SAD of
// input - ((p0 + p1 + 1) >> 1), with the SATD tiling and normalisation of
// Xg266ComputeAvgSatd (8x8 Hadamard when both sides are multiples of 8, else
// 4x4) -- for every block of a frame in one launch.  The plan flattens the
// blocks into their Hadamard tiles (8x8 when both sides are multiples of 8,
// else 4x4; SAD is tile-size independent, so the same tiles serve it): one
// thread per tile, so a frame of small blocks still fills the chip.  The plan
// keeps each block's tiles inside one workgroup of 256 (a block that would
// straddle two starts the next one, after empty tiles), so the workgroup sums
// its blocks in LDS and one thread per block writes the total: one launch,
// no pass that zeroes the outputs and no global atomics.
#include <algorithm>
#include <mutex>
#include <vector>

#include "xg_batch.hpp"
#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

struct BiPic {
    const uint16_t *data;
    intptr_t stride;
};

template <int N>
__device__ uint32_t bi_tile(const uint16_t *in, intptr_t is, const uint16_t *p0, intptr_t s0, const uint16_t *p1,
                            intptr_t s1, bool satd)
{
    int d[N * N];
#pragma unroll
    for (int y = 0; y < N; y++)
#pragma unroll
        for (int x = 0; x < N; x++)
            d[y * N + x] = (int)in[y * is + x] - (((int)p0[y * s0 + x] + (int)p1[y * s1 + x] + 1) >> 1);
    uint32_t s = 0;
    if (!satd) {
#pragma unroll
        for (int i = 0; i < N * N; i++) s += (uint32_t)abs(d[i]);
        return s;
    }
#pragma unroll
    for (int y = 0; y < N; y++) wht_reg<N>(d + y * N);
#pragma unroll
    for (int x = 0; x < N; x++) {
        int col[N];
#pragma unroll
        for (int y = 0; y < N; y++) col[y] = d[y * N + x];
        wht_reg<N>(col);
#pragma unroll
        for (int y = 0; y < N; y++) d[y * N + x] = col[y];
    }
#pragma unroll
    for (int i = 0; i < N * N; i++) s += (uint32_t)abs(d[i]);
    return N == 8 ? (s + 2) >> 2 : (s + 1) >> 1;
}

// one tile: block index, position (luma samples) and size (4 or 8)
struct BiTile {
    int32_t block;
    int16_t x, y;
    int32_t n;
};

constexpr int kBiWg = 256;                       // tiles per workgroup

__global__ void __launch_bounds__(kBiWg) k_bipred_tiles(BiPic in, BiPic p0, BiPic p1,
                                                        const BiTile *__restrict__ tiles, int nTiles, int satd,
                                                        uint32_t *__restrict__ out)
{
    __shared__ uint32_t sum[kBiWg];                  // the workgroup's blocks, from its first tile's
    const int i = (int)blockIdx.x * kBiWg + (int)threadIdx.x;
    const BiTile t = i < nTiles ? tiles[i] : BiTile{-1, 0, 0, 0};
    const int first = tiles[(int)blockIdx.x * kBiWg].block;   // a workgroup never starts with an empty tile
    sum[threadIdx.x] = 0;
    __syncthreads();
    if (t.block >= 0) {
        const uint16_t *a = in.data + (intptr_t)t.y * in.stride + t.x;
        const uint16_t *q0 = p0.data + (intptr_t)t.y * p0.stride + t.x;
        const uint16_t *q1 = p1.data + (intptr_t)t.y * p1.stride + t.x;
        const uint32_t c = t.n == 8 ? bi_tile<8>(a, in.stride, q0, p0.stride, q1, p1.stride, satd != 0)
                                    : bi_tile<4>(a, in.stride, q0, p0.stride, q1, p1.stride, satd != 0);
        atomicAdd(&sum[t.block - first], c);
    }
    __syncthreads();
    // the block's first tile writes its total
    if (t.block >= 0 && (threadIdx.x == 0 || tiles[i - 1].block != t.block)) out[t.block] = sum[t.block - first];
}

}  // namespace xg

using namespace xg;

namespace {
struct BiPlan {
    BiTile *dTiles = nullptr;
    int n = 0, nTiles = 0, cap = 0;
    int maxX = 0, maxY = 0;                     // extent of the blocks: the planes must cover it
    int widest = -1, tallest = -1;              // a block reaching maxX / maxY (for the error text)
    bool valid = false;
};
std::mutex g_bplan_mu;
std::vector<std::pair<xg_session *, BiPlan>> g_bplans;

BiPlan *bplan_for(xg_session *s, bool create)
{
    for (auto &p : g_bplans)
        if (p.first == s) return &p.second;
    if (!create) return nullptr;
    g_bplans.push_back({s, BiPlan()});
    return &g_bplans.back().second;
}
}  // namespace

extern "C" {

int32_t xg_bipred_plan(xg_session *s, const xg_block *blocks, int32_t n)
{
    if (!s || n < 0 || (n > 0 && !blocks)) {
        set_error(XG_ERR_ARG, "xg_bipred_plan: bad arguments");
        return XG_ERR_ARG;
    }
    for (int i = 0; i < n; i++) {
        const xg_block &b = blocks[i];
        if (b.w < 4 || b.h < 4 || b.w > 128 || b.h > 128 || (b.w & 3) || (b.h & 3) || b.x < 0 || b.y < 0) {
            set_error(XG_ERR_ARG, "xg_bipred_plan: block %d: sides must be multiples of 4 in 4..128 at x, y >= 0", i);
            std::lock_guard<std::mutex> g(g_bplan_mu);
            if (BiPlan *p = bplan_for(s, false)) p->valid = false;
            return XG_ERR_ARG;
        }
    }
    std::vector<BiTile> tiles;
    for (int i = 0; i < n; i++) {
        const xg_block &b = blocks[i];
        const int t = ((b.w & 7) == 0 && (b.h & 7) == 0) ? 8 : 4;
        const int k = (b.w / t) * (b.h / t);           // at most 256 (128x128 in 8x8 tiles)
        const int used = (int)tiles.size() % kBiWg;
        if (used && used + k > kBiWg) tiles.resize(tiles.size() + (kBiWg - used), BiTile{-1, 0, 0, 0});
        for (int y = 0; y < b.h; y += t)
            for (int x = 0; x < b.w; x += t) tiles.push_back(BiTile{i, (int16_t)(b.x + x), (int16_t)(b.y + y), t});
    }
    const int nt = (int)tiles.size();
    std::lock_guard<std::mutex> g(g_bplan_mu);
    BiPlan *p = bplan_for(s, true);
    p->valid = false;
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    if (!grow(p->dTiles, p->cap, std::max(nt, 1), st)) return XG_ERR_NOMEM;
    if (nt && !XG_HIP(hipMemcpyAsync(p->dTiles, tiles.data(), sizeof(BiTile) * (size_t)nt, hipMemcpyHostToDevice, st)))
        return XG_ERR_HIP;
    if (!XG_HIP(hipStreamSynchronize(st))) return XG_ERR_HIP;
    p->maxX = p->maxY = 0;
    p->widest = p->tallest = -1;
    for (int i = 0; i < n; i++) {
        if (blocks[i].x + blocks[i].w > p->maxX) p->maxX = blocks[i].x + blocks[i].w, p->widest = i;
        if (blocks[i].y + blocks[i].h > p->maxY) p->maxY = blocks[i].y + blocks[i].h, p->tallest = i;
    }
    p->n = n;
    p->nTiles = nt;
    p->valid = true;
    return XG_OK;
}

int32_t xg_bipred_run(xg_session *s, const xg_plane *input, const xg_plane *pred0, const xg_plane *pred1,
                      int32_t satd, uint32_t *dist)
{
    if (!s || !input || !pred0 || !pred1 || !input->data || !pred0->data || !pred1->data || !dist) {
        set_error(XG_ERR_ARG, "xg_bipred_run: bad arguments");
        return XG_ERR_ARG;
    }
    BiPlan *p;
    {
        std::lock_guard<std::mutex> g(g_bplan_mu);
        p = bplan_for(s, false);
    }
    if (!p || !p->valid) {
        set_error(XG_ERR_STATE, "xg_bipred_run: no plan (call xg_bipred_plan first; a failed plan leaves none)");
        return XG_ERR_STATE;
    }
    // every block inside all three planes (the plan keeps the blocks' extent)
    for (const xg_plane *pl : {input, pred0, pred1})
        if (p->maxX > pl->width || p->maxY > pl->height) {
            set_error(XG_ERR_ARG, "xg_bipred_run: block %d is outside a %dx%d plane",
                      p->maxX > pl->width ? p->widest : p->tallest, pl->width, pl->height);
            return XG_ERR_ARG;
        }
    if (p->n == 0) return XG_OK;
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    int ev[2];
    const bool timed = session_timer_begin(s, 7, ev);
    k_bipred_tiles<<<(p->nTiles + kBiWg - 1) / kBiWg, kBiWg, 0, st>>>(BiPic{input->data, input->stride},
                                                            BiPic{pred0->data, pred0->stride},
                                                            BiPic{pred1->data, pred1->stride}, p->dTiles, p->nTiles,
                                                            satd, dist);
    if (timed) session_timer_end(s, 7, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

}  // extern "C"

void xg::bipred_release(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_bplan_mu);
    for (size_t i = 0; i < g_bplans.size(); i++)
        if (g_bplans[i].first == s) {
            if (g_bplans[i].second.dTiles) (void)hipFree(g_bplans[i].second.dTiles);
            g_bplans.erase(g_bplans.begin() + (long)i);
            return;
        }
}

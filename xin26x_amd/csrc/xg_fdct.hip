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

#include "xg_device.hpp"
#include "xg_runtime.hpp"
#include "xg_trmatrix.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));

namespace xg {

// Padded int8 matrices: one per (type, lgN); rows padded to max(16, N),
// columns to max(32, N), zeros in the padding.
struct MatIndex { int off[3][7]; int cols[3][7]; int total; };

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
        if (!XG_HIP(hipMalloc((void **)&g_mats, h.size()))) return;
        if (!XG_HIP(hipMemcpy(g_mats, h.data(), h.size(), hipMemcpyHostToDevice))) return;
        g_mat_ok = true;
    });
    return g_mat_ok;
}

constexpr int kGroupMax = 4;

struct FdctGroup {
    uint32_t desc;                // lgW | lgH << 4 | horTx << 8 | verTx << 12
    int32_t count;                // blocks in the group, 1..kGroupMax
    long long off[kGroupMax];     // element offsets of the blocks in resid / coeff
};

// Blocks per group for a descriptor: enough to fill the 16-wide N dimension of
// both stages.
__host__ __device__ inline int group_cap(int H, int Wk)
{
    const int m = H < Wk ? H : Wk;
    const int c = m >= 16 ? 1 : 16 / m;
    return c < kGroupMax ? c : kGroupMax;
}

struct FdctArgs {
    const int16_t *resid;
    int32_t *coeff;
    const FdctGroup *groups;
    int n;
    const int8_t *mats;
    int matOff[3][7];
    int matCols[3][7];
};


__device__ __forceinline__ long pack8(const int *v)
{
    unsigned long long r = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) r |= (unsigned long long)(uint8_t)(int8_t)v[j] << (8 * j);
    return (long)r;
}

__device__ __forceinline__ long load_a(const int8_t *p)
{
    return *(const long *)p;      // 8-byte aligned: rows are multiples of 32, k offsets of 8
}

// HMAX: the largest block height of the launch; the LDS tile holds TMP[b][k][y]
// with HMAX + 4 int32 per (b, k) row, so launches of blocks up to 32 high hold
// half the LDS and run twice the workgroups per CU.
template <int HMAX>
__global__ void __launch_bounds__(64) k_fdct_mfma(FdctArgs a)
{
    __shared__ int32_t tmp[32 * (HMAX + 4)];
    const int lane = threadIdx.x;
    const int lr = lane & 15, lq = lane >> 4;
    for (int gi = blockIdx.x; gi < a.n; gi += gridDim.x) {
        const FdctGroup &G = a.groups[gi];
        const uint32_t desc = G.desc;
        const int G_n = G.count;
        const int lgW = desc & 15, lgH = (desc >> 4) & 15;
        const int hor = (desc >> 8) & 15, ver = (desc >> 12) & 15;
        const int W = 1 << lgW, H = 1 << lgH;
        const int Wk = tr_kept(hor, W), Hk = tr_kept(ver, H);
        const int lgWk = 31 - __builtin_clz(Wk);
        const int cap = group_cap(H, Wk);
        const int TS = H + 4;                       // LDS stride of a (b, k) row of TMP
        const int s1 = lgW + 1, s2 = lgH + 6;
        const int8_t *TW = a.mats + a.matOff[hor][lgW];
        const int8_t *TH = a.mats + a.matOff[ver][lgH];
        const int cw = a.matCols[hor][lgW], ch = a.matCols[ver][lgH];

        // ---- stage 1: TMP[b][k][y] over M = Wk, N = (b, y) rows of the group, K = W ----
        const int rows1 = cap * H;
        const int rnd1 = 1 << (s1 - 1);
        for (int m0 = 0; m0 < Wk; m0 += 16)
            for (int n0 = 0; n0 < rows1; n0 += 16) {
                v4i dlo = {0, 0, 0, 0}, dhi = {0, 0, 0, 0};
                const int row = n0 + lr, b = row >> lgH, y = row & (H - 1);
                const bool rowOk = b < G_n;
                const int16_t *R = a.resid + (rowOk ? G.off[b] : 0) + ((long)y << lgW);
                for (int k0 = 0; k0 < W; k0 += 32) {
                    const long av = load_a(TW + (m0 + lr) * cw + k0 + 8 * lq);
                    const int x = k0 + 8 * lq;
                    uint4 v = make_uint4(0, 0, 0, 0);
                    if (rowOk && x < W) {
                        if (W >= 8) {
                            v = *reinterpret_cast<const uint4 *>(R + x);
                        } else if (W == 4) {
                            const uint2 h2 = *reinterpret_cast<const uint2 *>(R + x);
                            v.x = h2.x;
                            v.y = h2.y;
                        } else {
                            v.x = *reinterpret_cast<const uint32_t *>(R + x);
                        }
                    }
                    // int8 limbs: lo = r & 127, hi = r >> 7, bytes in sample order
                    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
                    uint32_t m[4], hh[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        m[q] = w[q] & 0x007F007Fu;
                        hh[q] = as_u32(as_v2s(w[q]) >> (v2s){7, 7});
                    }
                    const long blo = (long)(((unsigned long long)__builtin_amdgcn_perm(m[3], m[2], 0x06040200u) << 32) |
                                            __builtin_amdgcn_perm(m[1], m[0], 0x06040200u));
                    const long bhi = (long)(((unsigned long long)__builtin_amdgcn_perm(hh[3], hh[2], 0x06040200u) << 32) |
                                            __builtin_amdgcn_perm(hh[1], hh[0], 0x06040200u));
                    dlo = __builtin_amdgcn_mfma_i32_16x16x32_i8(av, blo, dlo, 0, 0, 0);
                    dhi = __builtin_amdgcn_mfma_i32_16x16x32_i8(av, bhi, dhi, 0, 0, 0);
                }
                if (rowOk) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int k = m0 + 4 * lq + r;
                        if (k < Wk) tmp[((b << lgWk) + k) * TS + y] = (dlo[r] + (dhi[r] << 7) + rnd1) >> s1;
                    }
                }
            }
        __syncthreads();

        // ---- stage 2: OUT[b][j][k] over M = Hk, N = (b, k) columns of the group, K = H ----
        const int cols2 = cap * Wk;
        const int rnd2 = 1 << (s2 - 1);
        for (int m0 = 0; m0 < Hk; m0 += 16)
            for (int n0 = 0; n0 < cols2; n0 += 16) {
                v4i d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0}, d2 = {0, 0, 0, 0};
                const int col = n0 + lr, b = col >> lgWk, k = col & (Wk - 1);
                const bool colOk = b < G_n;
                const int32_t *trow = tmp + ((b << lgWk) + k) * TS;
                for (int k0 = 0; k0 < H; k0 += 32) {
                    const long av = load_a(TH + (m0 + lr) * ch + k0 + 8 * lq);
                    int l0[8], l1[8], l2[8];
                    const int yb = k0 + 8 * lq;
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const int t = (colOk && yb + j < H) ? trow[yb + j] : 0;
                        l0[j] = t & 127;
                        l1[j] = (t >> 7) & 127;
                        l2[j] = t >> 14;
                    }
                    d0 = __builtin_amdgcn_mfma_i32_16x16x32_i8(av, pack8(l0), d0, 0, 0, 0);
                    d1 = __builtin_amdgcn_mfma_i32_16x16x32_i8(av, pack8(l1), d1, 0, 0, 0);
                    d2 = __builtin_amdgcn_mfma_i32_16x16x32_i8(av, pack8(l2), d2, 0, 0, 0);
                }
                if (colOk) {
                    int32_t *O = a.coeff + G.off[b];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int j = m0 + 4 * lq + r;
                        if (j < Hk) O[j * W + k] = (d0[r] + (d1[r] << 7) + (d2[r] << 14) + rnd2) >> s2;
                    }
                }
            }
        // zero-out region (16-byte stores; kept widths are multiples of 4)
        if (Wk < W || Hk < H)
            for (int b = 0; b < G_n; b++) {
                int32_t *O = a.coeff + G.off[b];
                const int q4 = W >> 2;
                for (int idx = lane; idx < (W * H) >> 2; idx += 64) {
                    const int j = idx / q4, k = (idx - j * q4) << 2;
                    if (j >= Hk || k >= Wk) *reinterpret_cast<int4 *>(O + j * W + k) = make_int4(0, 0, 0, 0);
                }
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
    int n = 0, nLow = 0, cap = 0;     // groups
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
}  // namespace

extern "C" {

int32_t xg_fdct_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, int32_t n)
{
    if (!s || n < 0 || (n > 0 && (!offset || !desc))) {
        set_error(XG_ERR_ARG, "xg_fdct_plan: bad arguments");
        return XG_ERR_ARG;
    }
    if (!mats_init()) return XG_ERR_HIP;
    static const int8_t mts[6][2] = {{kDCT2, kDCT2}, {kSKIP, kSKIP}, {kDST7, kDST7},
                                     {kDCT8, kDST7}, {kDST7, kDCT8}, {kDCT8, kDCT8}};
    struct Item {
        long long off;
        uint32_t desc;
    };
    std::vector<Item> items((size_t)n);
    for (int i = 0; i < n; i++) {
        const int lgW = desc[i] & 15, lgH = (desc[i] >> 4) & 15, m = (desc[i] >> 8) & 15;
        if (lgW < 1 || lgW > 6 || lgH < 1 || lgH > 6 || m > 5 || m == 1) {
            set_error(XG_ERR_ARG, "xg_fdct_plan: block %d has descriptor 0x%x", i, desc[i]);
            return XG_ERR_ARG;
        }
        const int hor = mts[m][0], ver = mts[m][1];
        if ((hor != kDCT2 && (lgW < 2 || lgW > 5)) || (ver != kDCT2 && (lgH < 2 || lgH > 5)) || offset[i] < 0) {
            set_error(XG_ERR_ARG, "xg_fdct_plan: block %d: DST-VII/DCT-VIII need sides 4..32", i);
            return XG_ERR_ARG;
        }
        if (offset[i] % std::min(8, 1 << lgW)) {
            set_error(XG_ERR_ARG, "xg_fdct_plan: block %d: offset must be a multiple of min(8, width) samples", i);
            return XG_ERR_ARG;
        }
        items[i].off = offset[i];
        items[i].desc = (uint32_t)(lgW | (lgH << 4) | (hor << 8) | (ver << 12));
    }
    // two launches: blocks up to 32 high (small LDS tile), then 64-high blocks;
    // inside each, descriptors with the largest blocks first, then grouped
    std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
        const int hx = ((x.desc >> 4) & 15) == 6, hy = ((y.desc >> 4) & 15) == 6;
        if (hx != hy) return hx < hy;
        const int ax = (x.desc & 15) + ((x.desc >> 4) & 15), ay = (y.desc & 15) + ((y.desc >> 4) & 15);
        if (ax != ay) return ax > ay;
        return x.desc < y.desc;
    });
    std::vector<FdctGroup> groups;
    int nLow = 0;
    for (int i = 0; i < n;) {
        const uint32_t d = items[i].desc;
        const int lgW = d & 15, lgH = (d >> 4) & 15;
        const int cap = group_cap(1 << lgH, tr_kept((d >> 8) & 15, 1 << lgW));
        FdctGroup g;
        memset(&g, 0, sizeof(g));
        g.desc = d;
        while (i < n && items[i].desc == d && g.count < cap) g.off[g.count++] = items[i++].off;
        groups.push_back(g);
        if (lgH != 6) nLow = (int)groups.size();
    }
    const int ng = (int)groups.size();
    std::lock_guard<std::mutex> g(g_plan_mu);
    FdctPlan *p = plan_for(s, true);
    if (ng > p->cap) {
        if (p->dGroups) {
            hipStreamSynchronize((hipStream_t)xg_session_stream(s));
            hipFree(p->dGroups);
        }
        p->dGroups = nullptr;
        p->cap = 0;
        if (!XG_HIP(hipMalloc((void **)&p->dGroups, sizeof(FdctGroup) * (size_t)std::max(ng, 1)))) return XG_ERR_NOMEM;
        p->cap = std::max(ng, 1);
    }
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    if (ng && !XG_HIP(hipMemcpyAsync(p->dGroups, groups.data(), sizeof(FdctGroup) * (size_t)ng, hipMemcpyHostToDevice,
                                     st)))
        return XG_ERR_HIP;
    if (!XG_HIP(hipStreamSynchronize(st))) return XG_ERR_HIP;
    n = ng;
    p->n = n;
    p->nLow = nLow;
    return XG_OK;
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
    if (!p) {
        set_error(XG_ERR_STATE, "xg_fdct_run: no plan (call xg_fdct_plan first)");
        return XG_ERR_STATE;
    }
    if (p->n == 0) return XG_OK;
    FdctArgs a;
    a.resid = resid;
    a.coeff = coeff;
    a.groups = p->dGroups;
    a.n = p->n;
    a.mats = g_mats;
    memcpy(a.matOff, g_mi.off, sizeof(a.matOff));
    memcpy(a.matCols, g_mi.cols, sizeof(a.matCols));
    static int cus = [] {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            return prop.multiProcessorCount;
        return 256;
    }();
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    int ev[2];
    const bool timed = session_timer_begin(s, 2, ev);
    if (p->nLow > 0) {
        FdctArgs lo = a;
        lo.n = p->nLow;
        k_fdct_mfma<32><<<std::min(lo.n, cus * 32), 64, 0, st>>>(lo);
    }
    if (p->n > p->nLow) {
        FdctArgs hi = a;
        hi.groups = p->dGroups + p->nLow;
        hi.n = p->n - p->nLow;
        k_fdct_mfma<64><<<std::min(hi.n, cus * 16), 64, 0, st>>>(hi);
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
            if (g_plans[i].second.dGroups) hipFree(g_plans[i].second.dGroups);
            g_plans.erase(g_plans.begin() + (long)i);
            return;
        }
}

// xg_quant.hip -- batched quantisation + dequantisation of a batch of TUs.
//
// Per TU, pfXinQuantInvQuant as Xin266Transform calls it
// (h266_encoder/source/h266_trans_quant.c:204-222): over the
// min(W, 32) x min(H, 32) region, in coefficient groups (CGs) of
// cgW x cgH,
//   q = Clip16(sign(t) * ((|t| * qMult + qAdd) >> qShift))
//   r = Clip16((q * iqMult + iqAdd) >> iqShift)
// with int32 (wrapping) arithmetic as the reference's SINT32 expressions
// (h266_encoder/source/h266_quant_inv_quant.c:27-78), and bit k of the TU's
// map set when CG k (raster order in the region) holds a non-zero level.  The
// parameters come from the TU's descriptor as Xin266Transform derives them
// (h266_trans_quant.c:104-120; tables of h266_encoder_create.c:94-104,
// 1677-1705, 10-bit).  Past the region both outputs are zero (the reference
// zeroes the dequantised coefficients there, h266_trans_quant.c:163-188).
//
// Mapping: one thread per CG of each 64-item run.  The plan lists every TU's
// region CGs in raster order, then its CGs past the region, and pads so a
// TU's region CGs never straddle a run of 64: its map is then one ballot,
// shifted down to the lane of its first CG.  The kernel is HBM-bound elementwise work (4 bytes in,
// 2 + 2 out per coefficient); rows of 4 coefficients move as 16-byte loads and
// 8-byte stores.
#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "xg_batch.hpp"
#include "xg_device.hpp"
#include "xg_quant.hpp"
#include "xg_runtime.hpp"

namespace xg {

constexpr uint32_t kNoItem = 0xFFFFFFFFu;

// One CG: cgh rows of cgw coefficients (stride w), V per vector access.
template <int V>
__device__ __forceinline__ bool quant_cg(const int32_t *t, int16_t *q, int16_t *r, int w, int cgw, int cgh,
                                         const QParams &p)
{
    bool nz = false;
    for (int y = 0; y < cgh; y++)
        for (int x = 0; x < cgw; x += V) {
            const intptr_t o = (intptr_t)y * w + x;
            int tv[V], qv[V], rv[V];
            if constexpr (V == 4) {
                const int4 v = *reinterpret_cast<const int4 *>(t + o);
                tv[0] = v.x; tv[1] = v.y; tv[2] = v.z; tv[3] = v.w;
            } else if constexpr (V == 2) {
                const int2 v = *reinterpret_cast<const int2 *>(t + o);
                tv[0] = v.x; tv[1] = v.y;
            } else {
                tv[0] = t[o];
            }
#pragma unroll
            for (int j = 0; j < V; j++) {
                qv[j] = qiq(tv[j], p, rv[j]);
                nz |= qv[j] != 0;
            }
            if constexpr (V == 4) {
                *reinterpret_cast<uint2 *>(q + o) = make_uint2(pack16(qv[0], qv[1]), pack16(qv[2], qv[3]));
                *reinterpret_cast<uint2 *>(r + o) = make_uint2(pack16(rv[0], rv[1]), pack16(rv[2], rv[3]));
            } else if constexpr (V == 2) {
                *reinterpret_cast<uint32_t *>(q + o) = pack16(qv[0], qv[1]);
                *reinterpret_cast<uint32_t *>(r + o) = pack16(rv[0], rv[1]);
            } else {
                q[o] = (int16_t)qv[0];
                r[o] = (int16_t)rv[0];
            }
        }
    return nz;
}

template <int V>
__device__ __forceinline__ void zero_cg(int16_t *q, int16_t *r, int w, int cgw, int cgh)
{
    for (int y = 0; y < cgh; y++)
        for (int x = 0; x < cgw; x += V) {
            const intptr_t o = (intptr_t)y * w + x;
            if constexpr (V == 4) {
                *reinterpret_cast<uint2 *>(q + o) = make_uint2(0, 0);
                *reinterpret_cast<uint2 *>(r + o) = make_uint2(0, 0);
            } else if constexpr (V == 2) {
                *reinterpret_cast<uint32_t *>(q + o) = 0;
                *reinterpret_cast<uint32_t *>(r + o) = 0;
            } else {
                q[o] = 0;
                r[o] = 0;
            }
        }
}

struct QuantArgs {
    const int32_t *t;
    int16_t *q, *r;
    unsigned long long *nz;
    const uint32_t *items;         // tu << 8 | CG index (raster over the whole TU); kNoItem pads
    const int64_t *offset;
    const uint32_t *desc;
    int nItems;
};

// Items per lane: a wavefront takes kQSteps consecutive runs of 64 items (each
// run TU-aligned for the ballot) and issues every run's loads before the
// arithmetic, so each wave keeps kQSteps times the bytes in flight.
constexpr int kQSteps = 4;

__global__ void __launch_bounds__(256) k_quant(QuantArgs a)
{
    const int lane = threadIdx.x & 63;
    const int64_t wbase = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 * kQSteps);
    const int32_t *__restrict__ tc = a.t;
    int16_t *__restrict__ qc = a.q;
    int16_t *__restrict__ rc = a.r;
    uint32_t it[kQSteps], d[kQSteps];
    intptr_t o[kQSteps];
#pragma unroll
    for (int j = 0; j < kQSteps; j++) {
        const int64_t i = wbase + j * 64 + lane;
        it[j] = i < a.nItems ? a.items[i] : kNoItem;
    }
#pragma unroll
    for (int j = 0; j < kQSteps; j++) {
        d[j] = 0;
        o[j] = 0;
        if (it[j] != kNoItem) {
            const uint32_t tu = it[j] >> 8;
            d[j] = a.desc[tu];
            o[j] = (intptr_t)a.offset[tu];
        }
    }
    // a 4x4 group inside the region: its rows, loaded for every step up front
    int4 v[kQSteps][4];
    bool fast[kQSteps];
#pragma unroll
    for (int j = 0; j < kQSteps; j++) {
        const uint32_t dj = d[j];
        const int lgW = dj & 15, lgCW = (dj >> 12) & 7, lgCH = (dj >> 15) & 7;
        const int k = it[j] & 255, w = 1 << lgW;
        const int cx = k & ((1 << (lgW - lgCW)) - 1), cy = k >> (lgW - lgCW);
        const int x0 = cx << lgCW, y0 = cy << lgCH;
        o[j] += (intptr_t)y0 * w + x0;
        fast[j] = it[j] != kNoItem && lgCW == 2 && lgCH == 2 && x0 < 32 && y0 < 32;
        if (fast[j]) {
#pragma unroll
            for (int y = 0; y < 4; y++) v[j][y] = *reinterpret_cast<const int4 *>(tc + o[j] + (intptr_t)y * w);
        }
    }
#pragma unroll
    for (int j = 0; j < kQSteps; j++) {
        const uint32_t dj = d[j];
        const int lgW = dj & 15, lgH = (dj >> 4) & 15, lgCW = (dj >> 12) & 7, lgCH = (dj >> 15) & 7;
        const int k = it[j] & 255, w = 1 << lgW, cgw = 1 << lgCW, cgh = 1 << lgCH;
        const int cx = k & ((1 << (lgW - lgCW)) - 1), cy = k >> (lgW - lgCW);
        const int x0 = cx << lgCW, y0 = cy << lgCH;
        const bool inReg = it[j] != kNoItem && x0 < 32 && y0 < 32;
        bool nz = false;
        if (fast[j]) {
            const QParams p = quant_params(dj);
#pragma unroll
            for (int y = 0; y < 4; y++) {
                int qv[4], rv[4];
                qv[0] = qiq(v[j][y].x, p, rv[0]);
                qv[1] = qiq(v[j][y].y, p, rv[1]);
                qv[2] = qiq(v[j][y].z, p, rv[2]);
                qv[3] = qiq(v[j][y].w, p, rv[3]);
                nz |= (qv[0] | qv[1] | qv[2] | qv[3]) != 0;
                const intptr_t oy = o[j] + (intptr_t)y * w;
                *reinterpret_cast<uint2 *>(qc + oy) = make_uint2(pack16(qv[0], qv[1]), pack16(qv[2], qv[3]));
                *reinterpret_cast<uint2 *>(rc + oy) = make_uint2(pack16(rv[0], rv[1]), pack16(rv[2], rv[3]));
            }
        } else if (inReg) {
            const QParams p = quant_params(dj);
            nz = cgw >= 4 ? quant_cg<4>(tc + o[j], qc + o[j], rc + o[j], w, cgw, cgh, p)
               : cgw == 2 ? quant_cg<2>(tc + o[j], qc + o[j], rc + o[j], w, cgw, cgh, p)
                          : quant_cg<1>(tc + o[j], qc + o[j], rc + o[j], w, cgw, cgh, p);
        } else if (it[j] != kNoItem) {
            if (cgw >= 4) zero_cg<4>(qc + o[j], rc + o[j], w, cgw, cgh);
            else if (cgw == 2) zero_cg<2>(qc + o[j], rc + o[j], w, cgw, cgh);
            else zero_cg<1>(qc + o[j], rc + o[j], w, cgw, cgh);
        }
        const unsigned long long m = __ballot(nz);
        if (inReg) {
            const int rcx = min(w, 32) >> lgCW;
            if (cy * rcx + cx == 0) {
                const int nreg = rcx * (min(1 << lgH, 32) >> lgCH);
                a.nz[it[j] >> 8] = (m >> lane) & (nreg >= 64 ? ~0ull : ((1ull << nreg) - 1));
            }
        }
    }
}

}  // namespace xg

using namespace xg;

namespace {
struct QuantPlan {
    uint32_t *dItems = nullptr;
    int64_t *dOff = nullptr;
    uint32_t *dDesc = nullptr;
    int itemCap = 0, tuCap = 0, nItems = 0, n = 0;
    bool valid = false;
};
std::mutex g_qplan_mu;
std::vector<std::pair<xg_session *, QuantPlan>> g_qplans;

QuantPlan *qplan_for(xg_session *s, bool create)
{
    for (auto &p : g_qplans)
        if (p.first == s) return &p.second;
    if (!create) return nullptr;
    g_qplans.push_back({s, QuantPlan()});
    return &g_qplans.back().second;
}

int32_t qplan_fail(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_qplan_mu);
    if (QuantPlan *p = qplan_for(s, false)) p->valid = false;
    return XG_ERR_ARG;
}
}  // namespace

extern "C" {

int32_t xg_quant_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, int32_t n)
{
    if (!s || n < 0 || n >= (1 << 24) || (n > 0 && (!offset || !desc))) {
        set_error(XG_ERR_ARG, "xg_quant_plan: bad arguments (fewer than 2^24 TUs)");
        return qplan_fail(s);
    }
    std::vector<uint32_t> items;
    for (int i = 0; i < n; i++) {
        const uint32_t d = desc[i];
        const int lgW = d & 15, lgH = (d >> 4) & 15, mts = (d >> 8) & 15, lgCW = (d >> 12) & 7,
                  lgCH = (d >> 15) & 7, qp = (d >> 18) & 63;
        const int rw = std::min(1 << lgW, 32), rh = std::min(1 << lgH, 32);
        const int cgw = 1 << lgCW, cgh = 1 << lgCH;
        const int nreg = (rw / std::max(cgw, 1)) * (rh / std::max(cgh, 1));
        if (lgW < 1 || lgW > 6 || lgH < 1 || lgH > 6 || mts > 5 || (d >> 26) || cgw > rw || cgh > rh ||
            nreg > 64 || cgw * cgh > 64 || (lgW - lgCW) + (lgH - lgCH) > 8 || qp > 63 || offset[i] < 0 ||
            offset[i] % std::min(4, 1 << lgW)) {
            set_error(XG_ERR_ARG,
                      "xg_quant_plan: TU %d (descriptor 0x%x, offset %lld): sides 2..64, CGs inside the 32x32 "
                      "region (at most 64 there, 256 in the TU, 64 coefficients each), QP index <= 63, offset a "
                      "non-negative multiple of min(4, width)",
                      i, d, (long long)offset[i]);
            return qplan_fail(s);
        }
        const int ncx = (1 << lgW) >> lgCW, ncy = (1 << lgH) >> lgCH;
        const int rcx = rw >> lgCW, rcy = rh >> lgCH;
        if ((items.size() & 63) + (size_t)nreg > 64)   // keep the region CGs inside one wavefront
            items.resize((items.size() + 63) & ~(size_t)63, kNoItem);
        for (int cy = 0; cy < rcy; cy++)
            for (int cx = 0; cx < rcx; cx++) items.push_back((uint32_t)i << 8 | (uint32_t)(cy * ncx + cx));
        for (int cy = 0; cy < ncy; cy++)
            for (int cx = 0; cx < ncx; cx++)
                if (cx >= rcx || cy >= rcy) items.push_back((uint32_t)i << 8 | (uint32_t)(cy * ncx + cx));
    }
    if (items.size() > (size_t)INT32_MAX - 256) {
        set_error(XG_ERR_ARG, "xg_quant_plan: too many coefficient groups");
        return qplan_fail(s);
    }
    const int ni = (int)items.size();
    std::lock_guard<std::mutex> g(g_qplan_mu);
    QuantPlan *p = qplan_for(s, true);
    p->valid = false;
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    if (!grow(p->dItems, p->itemCap, std::max(ni, 1), st)) return XG_ERR_NOMEM;
    int tuCap = p->tuCap;
    if (!grow(p->dOff, tuCap, std::max(n, 1), st)) return XG_ERR_NOMEM;
    if (!grow(p->dDesc, p->tuCap, std::max(n, 1), st)) return XG_ERR_NOMEM;
    if (ni && !XG_HIP(hipMemcpyAsync(p->dItems, items.data(), sizeof(uint32_t) * (size_t)ni, hipMemcpyHostToDevice,
                                     st)))
        return XG_ERR_HIP;
    if (n && (!XG_HIP(hipMemcpyAsync(p->dOff, offset, sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, st)) ||
              !XG_HIP(hipMemcpyAsync(p->dDesc, desc, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, st))))
        return XG_ERR_HIP;
    if (!XG_HIP(hipStreamSynchronize(st))) return XG_ERR_HIP;
    p->nItems = ni;
    p->n = n;
    p->valid = true;
    return XG_OK;
}

int32_t xg_quant_run(xg_session *s, const xg_coeff *tCoeff, int16_t *qCoeff, int16_t *rCoeff, uint64_t *nzMap)
{
    if (!s || !tCoeff || !qCoeff || !rCoeff || !nzMap || ((uintptr_t)tCoeff & 15) || ((uintptr_t)qCoeff & 7) ||
        ((uintptr_t)rCoeff & 7) || ((uintptr_t)nzMap & 7)) {
        set_error(XG_ERR_ARG, "xg_quant_run: bad arguments (tCoeff 16-byte aligned, outputs 8-byte aligned)");
        return XG_ERR_ARG;
    }
    QuantPlan *p;
    {
        std::lock_guard<std::mutex> g(g_qplan_mu);
        p = qplan_for(s, false);
    }
    if (!p || !p->valid) {
        set_error(XG_ERR_STATE, "xg_quant_run: no plan (call xg_quant_plan first; a failed plan leaves none)");
        return XG_ERR_STATE;
    }
    if (p->nItems == 0) return XG_OK;
    QuantArgs a;
    a.t = tCoeff;
    a.q = qCoeff;
    a.r = rCoeff;
    a.nz = reinterpret_cast<unsigned long long *>(nzMap);
    a.items = p->dItems;
    a.offset = p->dOff;
    a.desc = p->dDesc;
    a.nItems = p->nItems;
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    int ev[2];
    const bool timed = session_timer_begin(s, 8, ev);
    k_quant<<<(int)((p->nItems + 256 * kQSteps - 1) / (256 * kQSteps)), 256, 0, st>>>(a);
    if (timed) session_timer_end(s, 8, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

}  // extern "C"

void xg::quant_release(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_qplan_mu);
    for (size_t i = 0; i < g_qplans.size(); i++)
        if (g_qplans[i].first == s) {
            QuantPlan &p = g_qplans[i].second;
            if (p.dItems) (void)hipFree(p.dItems);
            if (p.dOff) (void)hipFree(p.dOff);
            if (p.dDesc) (void)hipFree(p.dDesc);
            g_qplans.erase(g_qplans.begin() + (long)i);
            return;
        }
}

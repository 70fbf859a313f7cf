// xg_lookahead.hip -- lookahead motion estimation on half-resolution pictures.
//
// The reference's lookahead (Xin26xLookaheadFrame, h26x_core/include/h26x_look_ahead.h:27)
// searches 2x-downscaled pictures per unit of laUnitSize samples (8 with B frames, 16
// without, h266_encoder/source/h266_encoder_create.c:1745) and leaves one MV per unit in
// laMv; the inter CU search takes laMv << 1 as a full-resolution start candidate
// (h266_encoder/source/h266_analyze_inter_cu.c:247-261).  Its downscale
// (Xin26xDownscale2x2) and its search are closed, so this entry is a design of its own
// that fills the same role:
//   * downscale: each half-resolution sample is the rounded mean of its 2x2 source
//     block, (a + b + c + d + 2) >> 2.  The reference picture is downscaled over its
//     padding too (as h266_enc_init.c:2804-2813 downscales the padded buffer), so the
//     half-resolution reference carries half the padding;
//   * search: the integer full search of xg_full_search_frame on every unit wholly
//     inside the half-resolution picture, predictor (0, 0), SAD + lambda * MV bits.
#include <algorithm>
#include <mutex>
#include <vector>

#include "xg_device.hpp"
#include "xg_runtime.hpp"

namespace xg {

// dst (w x h) = rounded 2x2 means of src; src points at the top-left source sample
__global__ void k_down2x2(const uint16_t *__restrict__ src, intptr_t ss, uint16_t *__restrict__ dst, intptr_t ds,
                          int w, int h)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    const uint16_t *p = src + (intptr_t)(2 * y) * ss + 2 * x;
    dst[(intptr_t)y * ds + x] = (uint16_t)(((int)p[0] + p[1] + p[ss] + p[ss + 1] + 2) >> 2);
}

}  // namespace xg

using namespace xg;

namespace {
struct HalfPics {
    uint16_t *cur = nullptr, *ref = nullptr;
    size_t curCap = 0, refCap = 0;
};
std::mutex g_la_mu;
std::vector<std::pair<xg_session *, HalfPics>> g_la;

HalfPics *half_for(xg_session *s)
{
    for (auto &p : g_la)
        if (p.first == s) return &p.second;
    g_la.push_back({s, HalfPics()});
    return &g_la.back().second;
}

bool ensure(uint16_t *&p, size_t &cap, size_t n, hipStream_t st)
{
    if (n <= cap) return true;
    if (p) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(p);
    }
    p = nullptr;
    cap = 0;
    if (!XG_HIP(hipMalloc((void **)&p, n * sizeof(uint16_t)))) return false;
    cap = n;
    return true;
}

// Enqueues the two downscales into the half-resolution planes `hc` (current) and
// `hr` (reference, hp samples of padding) and the unit search on `st`.
int32_t lookahead_launch(hipStream_t st, xg_session *timer, const xg_plane *cur, const xg_plane *ref,
                         const xg_plane &hc, const xg_plane &hr, int32_t unitSize, int32_t range, uint32_t sadLambda,
                         int16_t *laMv, uint32_t *cost, uint32_t *sad)
{
    const int hp = hr.pad, rw = hr.stride, rh = hr.height + 2 * hp;
    k_down2x2<<<dim3((hc.width + 127) / 128, hc.height), 128, 0, st>>>(cur->data, cur->stride, hc.data, hc.stride,
                                                                       hc.width, hc.height);
    k_down2x2<<<dim3((rw + 127) / 128, rh), 128, 0, st>>>(ref->data - (intptr_t)ref->pad * ref->stride - ref->pad,
                                                           ref->stride, hr.data - (intptr_t)hp * rw - hp, rw, rw, rh);
    if (!XG_HIP(hipGetLastError())) return XG_ERR_HIP;
    // window of a CU with ctuSize = range: MVs up to range + 8 samples past the picture
    // edge, so the search keeps its full +/-range everywhere
    return full_search_frame_on(st, timer, &hc, &hr, unitSize, unitSize, range, sadLambda, range, laMv, cost, sad);
}

bool lookahead_args_ok(const char *fn, const xg_plane *cur, const xg_plane *ref, int32_t unitSize, int32_t range,
                       const int16_t *laMv, const uint32_t *cost, const uint32_t *sad)
{
    if (!cur || !ref || !cur->data || !ref->data || !laMv || !cost || !sad || (unitSize != 8 && unitSize != 16) ||
        range < 0 || range > 64 || cur->width != ref->width || cur->height != ref->height || cur->width < 2 ||
        cur->height < 2 || (ref->pad & 1) || ref->pad / 2 < range + 4) {
        set_error(XG_ERR_ARG, "%s: bad arguments (units 8 or 16, range <= 64, reference padding even "
                              "and >= 2 * (range + 4))", fn);
        return false;
    }
    return true;
}

// Samples of the half-resolution current plane (even stride) and reference
// plane (with its padding) for a cur->width x cur->height picture.
void half_sizes(const xg_plane *cur, const xg_plane *ref, size_t *curN, size_t *refN)
{
    const int hw = cur->width / 2, hh = cur->height / 2, hp = ref->pad / 2;
    *curN = (size_t)((hw + 1) & ~1) * hh;
    *refN = (size_t)(hw + 2 * hp) * (hh + 2 * hp);
}

}  // namespace

// ---- the no-allocation form ------------------------------------------------
// Both half-resolution planes of xg_lookahead_me_static live in this array:
// the current plane at sample 0, the reference plane at the next multiple of
// 128 samples (256 bytes).  Every device the library runs on holds its own
// copy for the life of the process.  Its host-side name is not a device
// address: host code uses only the per-device pointer from hipGetSymbolAddress.
static constexpr size_t kLaScratchSamples = (size_t)XG_LOOKAHEAD_STATIC_MAX_BYTES / sizeof(uint16_t);
static __device__ __align__(256) uint16_t scratch[kLaScratchSamples];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint16_t *g_scratch[kMaxDevices] = {};

// The calling thread's current device's copy of `scratch`, resolved once per
// device with that device current; nullptr (error recorded) on failure.
uint16_t *scratch_on_current_device()
{
    int dev = -1;
    if (!XG_HIP(hipGetDevice(&dev))) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error(XG_ERR_ARG, "xg_lookahead_me_static: device %d beyond the %d the library tracks", dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (!XG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) || !p) return nullptr;
        g_scratch[dev] = (uint16_t *)p;
    }
    return g_scratch[dev];
}
}  // namespace

extern "C" {

int32_t xg_lookahead_me(xg_session *s, const xg_plane *cur, const xg_plane *ref, int32_t unitSize, int32_t range,
                        uint32_t sadLambda, int16_t *laMv, uint32_t *cost, uint32_t *sad)
{
    if (!s) {
        set_error(XG_ERR_ARG, "xg_lookahead_me: bad arguments (no session)");
        return XG_ERR_ARG;
    }
    if (!lookahead_args_ok("xg_lookahead_me", cur, ref, unitSize, range, laMv, cost, sad)) return XG_ERR_ARG;
    const int hw = cur->width / 2, hh = cur->height / 2, hp = ref->pad / 2;
    const int rw = hw + 2 * hp;
    size_t curN, refN;
    half_sizes(cur, ref, &curN, &refN);
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    xg_plane hc, hr;
    {
        std::lock_guard<std::mutex> g(g_la_mu);
        HalfPics *h = half_for(s);
        // the current picture's stride is kept even, so the packed search kernel takes it
        const int cs = (hw + 1) & ~1;
        if (!ensure(h->cur, h->curCap, curN, st) || !ensure(h->ref, h->refCap, refN, st)) return XG_ERR_NOMEM;
        hc = xg_plane{h->cur, cs, hw, hh, 0};
        hr = xg_plane{h->ref + (size_t)hp * rw + hp, rw, hw, hh, hp};
    }
    return lookahead_launch(st, s, cur, ref, hc, hr, unitSize, range, sadLambda, laMv, cost, sad);
}

int32_t xg_lookahead_me_static(void *stream, const xg_plane *cur, const xg_plane *ref, int32_t unitSize,
                               int32_t range, uint32_t sadLambda, int16_t *laMv, uint32_t *cost, uint32_t *sad)
{
    if (!lookahead_args_ok("xg_lookahead_me_static", cur, ref, unitSize, range, laMv, cost, sad)) return XG_ERR_ARG;
    size_t curN, refN;
    half_sizes(cur, ref, &curN, &refN);
    const size_t refOff = align_up(curN, 128);
    if (refOff + refN > kLaScratchSamples) {
        set_error(XG_ERR_ARG, "xg_lookahead_me_static: the half-resolution pictures need %zu bytes, more than the "
                  "%zu of the static scratch (XG_LOOKAHEAD_STATIC_MAX_BYTES)", (refOff + refN) * sizeof(uint16_t),
                  kLaScratchSamples * sizeof(uint16_t));
        return XG_ERR_ARG;
    }
    uint16_t *base = scratch_on_current_device();
    if (!base) return XG_ERR_HIP;
    const int hw = cur->width / 2, hh = cur->height / 2, hp = ref->pad / 2;
    const int rw = hw + 2 * hp;
    const xg_plane hc{base, (hw + 1) & ~1, hw, hh, 0};
    const xg_plane hr{base + refOff + (size_t)hp * rw + hp, rw, hw, hh, hp};
    return lookahead_launch((hipStream_t)stream, nullptr, cur, ref, hc, hr, unitSize, range, sadLambda, laMv, cost,
                            sad);
}

}  // extern "C"

void xg::lookahead_release(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_la_mu);
    for (size_t i = 0; i < g_la.size(); i++)
        if (g_la[i].first == s) {
            if (g_la[i].second.cur) (void)hipFree(g_la[i].second.cur);
            if (g_la[i].second.ref) (void)hipFree(g_la[i].second.ref);
            g_la.erase(g_la.begin() + (long)i);
            return;
        }
}

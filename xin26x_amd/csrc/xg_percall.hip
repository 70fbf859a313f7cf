// xg_percall.hip -- the per-call primitives of the xin_func_struct mirror.
//
// Each entry point takes the reference's host pointers, stages exactly the
// samples the reference function reads into pinned memory (samples it does
// not read are zero-filled and only ever meet zero filter taps), copies them
// to HBM on the calling thread's stream, runs one kernel and copies the
// result back before returning.
//
// Interpolation: every table slot (Copy, Hor, Vet, HorVet and their S16
// variants) runs the same separable kernel with a phase-0 filter in the
// direction the slot does not filter.  For 10-bit samples that is
// bit-identical to the reference's separate code paths
// (h266_core/source/h266_luma_mc.c:72-559, h266_chroma_mc.c:62-384):
//   Hor:  clip((S+32)>>6)        == clip((64*((S-32768)>>2) + 524800) >> 10)
//   Vet:  clip((S+32)>>6)        with first-pass value 16*s - 8192 (exact)
//   Copy: s                      (both passes phase 0)
//   S16:  (S-32768)>>2 == (64*t)>>6, (16*S - 2^19)>>6 == (S-32768)>>2
// so the kernels have no per-slot branches.
#include <mutex>
#include <string.h>

#include "xg_device.hpp"
#include "xg_runtime.hpp"
#include "xg_trmatrix.hpp"

namespace xg {

static __constant__ TrTables c_tr = kTr;

static std::once_flag g_init_once;
static bool g_init_ok = false;

bool device_init()
{
    std::call_once(g_init_once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            set_error(XG_ERR_HIP, "no HIP device available");
            return;
        }
        hipDeviceProp_t prop;
        if (!XG_HIP(hipGetDeviceProperties(&prop, 0))) return;
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            log_msg(1, "device is %s; this library is built for gfx950", prop.gcnArchName);
        g_init_ok = true;
    });
    if (!g_init_ok) set_error(XG_ERR_HIP, "HIP device initialisation failed");
    return g_init_ok;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Separable interpolation of a packed region: src holds (h+TAPS-1) rows of
// (w+TAPS-1) samples starting TAPS/2-1 samples above/left of the block.
template <int TAPS, bool S16>
__global__ void k_interp(const uint16_t *__restrict__ src, int w, int h, int fx, int fy, int fi,
                         void *__restrict__ dst)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w || y >= h) return;
    const int ss = w + TAPS - 1;
    const int8_t *ch = taps_for<TAPS>(fx, fi);
    const int8_t *cv = taps_for<TAPS>(fy, fi);
    const uint16_t *p = src + y * ss + x;
    int acc = 0;
#pragma unroll
    for (int k = 0; k < TAPS; k++) acc += (int)cv[k] * hpass<TAPS>(p + k * ss, ch);
    if constexpr (S16) ((int16_t *)dst)[y * w + x] = (int16_t)vfinal_s16(acc);
    else ((uint16_t *)dst)[y * w + x] = (uint16_t)vfinal_pix(acc);
}

// Vertical pass over an S16 first-pass buffer (VetS16U8 / VetS16S16):
// src holds (h+TAPS-1) rows of w values starting TAPS/2-1 rows above.
template <int TAPS, bool S16>
__global__ void k_vet_s16(const int16_t *__restrict__ src, int w, int h, int fy, int fi,
                          void *__restrict__ dst)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w || y >= h) return;
    const int8_t *cv = taps_for<TAPS>(fy, fi);
    int acc = 0;
#pragma unroll
    for (int k = 0; k < TAPS; k++) acc += (int)cv[k] * (int)src[(y + k) * w + x];
    if constexpr (S16) ((int16_t *)dst)[y * w + x] = (int16_t)vfinal_s16(acc);
    else ((uint16_t *)dst)[y * w + x] = (uint16_t)vfinal_pix(acc);
}

// Bi-prediction average (h266_luma_mc.c:561-593) or BCW weighting (:595-629).
__global__ void k_bipred(const int16_t *__restrict__ a, const int16_t *__restrict__ b, int n,
                         int weighted, int wa, int wb, uint16_t *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v;
    if (weighted) v = ((int)a[i] * wa + (int)b[i] * wb + (1 << 6) + (8192 << 3)) >> 7;
    else v = ((int)a[i] + (int)b[i] + (1 << 4) + 2 * 8192) >> 5;
    dst[i] = (uint16_t)clip_pix(v);
}

// SAD of packed reference blocks (one every refPitch samples) against one
// packed input block; one workgroup per reference block.
__global__ void __launch_bounds__(256) k_sad(const uint16_t *__restrict__ in,
                                             const uint16_t *__restrict__ refs, int npix, int refPitch,
                                             uint32_t *__restrict__ out)
{
    const uint16_t *r = refs + (size_t)blockIdx.x * refPitch;
    uint32_t s = 0;
    for (int i = threadIdx.x; i < npix; i += 256) s += (uint32_t)abs((int)in[i] - (int)r[i]);
    s = block_sum256(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// Rounded average (a + b + 1) >> 1 of packed block pairs, in place into a:
// pair k at a + k * pitch, b + k * pitch (blockIdx.y = k).
__global__ void k_avg2(uint16_t *__restrict__ a, const uint16_t *__restrict__ b, int n, int pitch)
{
    const size_t base = (size_t)blockIdx.y * pitch;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        a[base + i] = (uint16_t)(((uint32_t)a[base + i] + b[base + i] + 1) >> 1);
}

// Quantisation + inverse quantisation of one packed w x h block
// (h266_encoder/source/h266_quant_inv_quant.c:27-143): one lane per coefficient
// group (at most 64), the non-zero map from a wave ballot, so group k (raster
// order) lands on bit k as in the reference's reverse walk with left shifts.
// Arithmetic is int32 as the reference's SINT32 expressions (wrapping).
__global__ void __launch_bounds__(64) k_qiq(const int32_t *__restrict__ t, int32_t *__restrict__ q,
                                            int32_t *__restrict__ r, int w, int h, int cgw, int cgh, int qMult,
                                            int qAdd, int qShift, int iqMult, int iqAdd, int iqShift,
                                            unsigned long long *__restrict__ map)
{
    const int ncx = w / cgw, ncg = ncx * (h / cgh), g = threadIdx.x;
    bool nz = false;
    if (g < ncg) {
        const int x0 = (g % ncx) * cgw, y0 = (g / ncx) * cgh;
        for (int y = y0; y < y0 + cgh; y++)
            for (int x = x0; x < x0 + cgw; x++) {
                const int i = y * w + x;
                const int tc = t[i], a = tc < 0 ? -tc : tc;
                int qc = (int)((unsigned)a * (unsigned)qMult + (unsigned)qAdd) >> qShift;
                qc = tc < 0 ? -qc : qc;
                qc = min(max(qc, -32768), 32767);
                q[i] = qc;
                const int rc = (int)((unsigned)qc * (unsigned)iqMult + (unsigned)iqAdd) >> iqShift;
                r[i] = min(max(rc, -32768), 32767);
                nz |= qc != 0;
            }
    }
    const unsigned long long m = __ballot(nz);
    if (g == 0) *map = m;
}

// SATD of a packed w x h block pair; one thread per Hadamard tile.
template <int N>
__device__ uint32_t tile_satd_serial(const uint16_t *a, const uint16_t *b, int stride)
{
    int d[N * N];
#pragma unroll
    for (int y = 0; y < N; y++)
#pragma unroll
        for (int x = 0; x < N; x++) d[y * N + x] = (int)a[y * stride + x] - (int)b[y * stride + x];
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
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < N * N; i++) s += (uint32_t)abs(d[i]);
    if (N == 8) return (s + 2) >> 2;
    if (N == 4) return (s + 1) >> 1;
    return s;
}

__global__ void __launch_bounds__(256) k_satd(const uint16_t *__restrict__ a,
                                              const uint16_t *__restrict__ b, int w, int h,
                                              uint32_t *__restrict__ out)
{
    const int n = (w % 8 == 0 && h % 8 == 0) ? 8 : (w % 4 == 0 && h % 4 == 0) ? 4 : 2;
    const int tw = w / n, tiles = tw * (h / n);
    uint32_t s = 0;
    for (int t = threadIdx.x; t < tiles; t += 256) {
        const int off = (t / tw) * n * w + (t % tw) * n;
        if (n == 8) s += tile_satd_serial<8>(a + off, b + off, w);
        else if (n == 4) s += tile_satd_serial<4>(a + off, b + off, w);
        else s += tile_satd_serial<2>(a + off, b + off, w);
    }
    s = block_sum256(s);
    if (threadIdx.x == 0) out[0] = s;
}

// 1-D forward transform: in holds `line` vectors of N values (packed),
// out[k*line + j] (packed).  int64 accumulation, round-half-up shift.
__global__ void k_fdct1d(const int32_t *__restrict__ in, int type, int N, int line, int shift,
                         int32_t *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= line || k >= N) return;
    int32_t v = 0;
    if (k < tr_kept(type, N)) {
        long long s = 0;
        for (int n = 0; n < N; n++) s += (long long)tr_entry(c_tr, type, N, k, n) * in[(size_t)j * N + n];
        const long long rnd = shift > 0 ? (1LL << (shift - 1)) : 0;
        v = (int32_t)((s + rnd) >> shift);
    }
    out[(size_t)k * line + j] = v;
}

// 1-D inverse transform (H.266 8.7.4.5 with the nonZeroS bound of 8.7.4.2):
// in packed [N][line] (coefficient k of line j at in[k*line + j]), out packed
// [line][N].  Lines j >= line - skip are written as zero; each output is
// rounded, shifted and clipped to int16.
__global__ void k_idct1d(const int32_t *__restrict__ in, int type, int N, int line, int skip, int shift,
                         int32_t *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= line || i >= N) return;
    int32_t v = 0;
    if (j < line - skip) {
        long long s = 0;
        const int nz = tr_kept(type, N);
        for (int k = 0; k < nz; k++) s += (long long)tr_entry(c_tr, type, N, k, i) * in[(size_t)k * line + j];
        const long long rnd = shift > 0 ? (1LL << (shift - 1)) : 0;
        const long long r = (s + rnd) >> shift;
        v = (int32_t)(r < -32768 ? -32768 : r > 32767 ? 32767 : r);
    }
    out[(size_t)j * N + i] = v;
}

// Transform-skip inverse: arithmetic shift, truncated to int16.
__global__ void k_iskip(const int32_t *__restrict__ in, int n, int shift, int32_t *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int16_t)(in[i] >> shift);
}

// Reconstruction: Clip3(0, 1023, pred + residual).
__global__ void k_recon(const int32_t *__restrict__ res, const uint16_t *__restrict__ pred, int n,
                        uint16_t *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = (int)pred[i] + res[i];
    out[i] = (uint16_t)(v < 0 ? 0 : v > 1023 ? 1023 : v);
}

__global__ void k_blocksub(const uint16_t *__restrict__ a, const uint16_t *__restrict__ b, int n,
                           int32_t *__restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (int32_t)a[i] - (int32_t)b[i];
}

// ---------------------------------------------------------------------------
// Host wrappers
// ---------------------------------------------------------------------------
enum Dir { kCopy = 0, kHor = 1, kVet = 2, kHorVet = 3 };

static bool check_wh(uint32_t w, uint32_t h, uint32_t maxw, uint32_t maxh, const char *fn)
{
    if (w == 0 || h == 0 || w > maxw || h > maxh) {
        set_error(XG_ERR_ARG, "%s: unsupported block %ux%u", fn, w, h);
        return false;
    }
    return true;
}

static bool finish(CallCtx *c) { return XG_HIP(hipStreamSynchronize(c->stream)); }

// PIXEL source, PIXEL or S16 destination.
static void interp_u16(int taps, bool outS16, Dir dir, const uint16_t *src, intptr_t ss, void *dst,
                       intptr_t ds, int fx, int fy, int fi, uint32_t w, uint32_t h, const char *fn)
{
    if (!check_wh(w, h, 128, 128, fn)) return;
    if (fi < 0 || fi > 2) fi = 0;   // the reference treats any other index as the default filter
    const int hb = taps / 2 - 1, rw = (int)w + taps - 1, rh = (int)h + taps - 1;
    const size_t inB = align_up((size_t)rw * rh * 2, 256), outB = (size_t)w * h * 2;
    CallCtx *c = call_ctx(inB + outB, inB + outB);
    if (!c) return;
    uint16_t *hin = (uint16_t *)c->host;
    memset(hin, 0, (size_t)rw * rh * 2);
    const bool rows = dir == kVet || dir == kHorVet, cols = dir == kHor || dir == kHorVet;
    const int r0 = rows ? 0 : hb, r1 = rows ? rh : hb + (int)h;
    const int c0 = cols ? 0 : hb, c1 = cols ? rw : hb + (int)w;
    for (int r = r0; r < r1; r++)
        memcpy(hin + (size_t)r * rw + c0, src + (intptr_t)(r - hb) * ss + (c0 - hb), (size_t)(c1 - c0) * 2);
    char *din = c->dev, *dout = c->dev + inB;
    if (!XG_HIP(hipMemcpyAsync(din, hin, (size_t)rw * rh * 2, hipMemcpyHostToDevice, c->stream))) return;
    dim3 grid((w + 63) / 64, h), block(64);
    if (taps == 8) {
        if (outS16) k_interp<8, true><<<grid, block, 0, c->stream>>>((uint16_t *)din, w, h, fx, fy, fi, dout);
        else k_interp<8, false><<<grid, block, 0, c->stream>>>((uint16_t *)din, w, h, fx, fy, fi, dout);
    } else {
        if (outS16) k_interp<4, true><<<grid, block, 0, c->stream>>>((uint16_t *)din, w, h, fx, fy, 0, dout);
        else k_interp<4, false><<<grid, block, 0, c->stream>>>((uint16_t *)din, w, h, fx, fy, 0, dout);
    }
    if (!XG_HIP(hipGetLastError())) return;
    char *hout = c->host + inB;
    if (!XG_HIP(hipMemcpyAsync(hout, dout, outB, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    unpack2d(dst, ds, hout, (int)w, (int)h, 2);
}

// S16 source (first-pass buffer), vertical pass only.
static void interp_s16(int taps, bool outS16, const int16_t *src, intptr_t ss, void *dst, intptr_t ds,
                       int fy, int fi, uint32_t w, uint32_t h, const char *fn)
{
    if (!check_wh(w, h, 128, 128, fn)) return;
    if (fi < 0 || fi > 2) fi = 0;   // the reference treats any other index as the default filter
    const int hb = taps / 2 - 1, rh = (int)h + taps - 1;
    const size_t inB = align_up((size_t)w * rh * 2, 256), outB = (size_t)w * h * 2;
    CallCtx *c = call_ctx(inB + outB, inB + outB);
    if (!c) return;
    pack2d(c->host, src - (intptr_t)hb * ss, ss, (int)w, rh, 2);
    char *din = c->dev, *dout = c->dev + inB;
    if (!XG_HIP(hipMemcpyAsync(din, c->host, (size_t)w * rh * 2, hipMemcpyHostToDevice, c->stream))) return;
    dim3 grid((w + 63) / 64, h), block(64);
    if (taps == 8) {
        if (outS16) k_vet_s16<8, true><<<grid, block, 0, c->stream>>>((int16_t *)din, w, h, fy, fi, dout);
        else k_vet_s16<8, false><<<grid, block, 0, c->stream>>>((int16_t *)din, w, h, fy, fi, dout);
    } else {
        if (outS16) k_vet_s16<4, true><<<grid, block, 0, c->stream>>>((int16_t *)din, w, h, fy, 0, dout);
        else k_vet_s16<4, false><<<grid, block, 0, c->stream>>>((int16_t *)din, w, h, fy, 0, dout);
    }
    if (!XG_HIP(hipGetLastError())) return;
    char *hout = c->host + inB;
    if (!XG_HIP(hipMemcpyAsync(hout, dout, outB, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    unpack2d(dst, ds, hout, (int)w, (int)h, 2);
}

static void bipred(const int16_t *a, intptr_t sa, const int16_t *b, intptr_t sb, uint16_t *dst,
                   intptr_t ds, uint32_t w, uint32_t h, int weighted, int wa, int wb, const char *fn)
{
    if (!check_wh(w, h, 128, 128, fn)) return;
    const size_t n = (size_t)w * h, blk = align_up(n * 2, 256);
    CallCtx *c = call_ctx(3 * blk, 3 * blk);
    if (!c) return;
    pack2d(c->host, a, sa, (int)w, (int)h, 2);
    pack2d(c->host + blk, b, sb, (int)w, (int)h, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, 2 * blk, hipMemcpyHostToDevice, c->stream))) return;
    k_bipred<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>((int16_t *)c->dev, (int16_t *)(c->dev + blk),
                                                               (int)n, weighted, wa, wb,
                                                               (uint16_t *)(c->dev + 2 * blk));
    if (!XG_HIP(hipGetLastError())) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + 2 * blk, c->dev + 2 * blk, n * 2, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    unpack2d(dst, ds, c->host + 2 * blk, (int)w, (int)h, 2);
}

// SAD of one input block against nref reference blocks.
static void sad_multi(const uint16_t *in, intptr_t is, const uint16_t *const *refs, int nref,
                      intptr_t rs, uint32_t w, uint32_t h, uint32_t *out, const char *fn)
{
    for (int k = 0; k < nref; k++) out[k] = UINT32_MAX;
    if (!check_wh(w, h, 128, 128, fn)) return;
    const size_t npix = (size_t)w * h, blk = align_up(npix * 2, 256);
    const size_t bytes = blk * (1 + nref) + 256;
    CallCtx *c = call_ctx(bytes, bytes);
    if (!c) return;
    pack2d(c->host, in, is, (int)w, (int)h, 2);
    for (int k = 0; k < nref; k++) pack2d(c->host + blk * (1 + k), refs[k], rs, (int)w, (int)h, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, blk * (1 + nref), hipMemcpyHostToDevice, c->stream))) return;
    uint32_t *dres = (uint32_t *)(c->dev + blk * (1 + nref));
    k_sad<<<nref, 256, 0, c->stream>>>((uint16_t *)c->dev, (uint16_t *)(c->dev + blk), (int)npix, (int)(blk / 2),
                                       dres);
    if (!XG_HIP(hipGetLastError())) return;
    uint32_t *hres = (uint32_t *)(c->host + blk * (1 + nref));
    if (!XG_HIP(hipMemcpyAsync(hres, dres, 4 * nref, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    for (int k = 0; k < nref; k++) out[k] = hres[k];
}

// Distortion of one input block against the rounded averages of nref
// prediction pairs (refA[k], refB[k]): SAD for each pair, or the SATD of the
// single pair (satd, nref == 1).
static void avg_dist(const uint16_t *in, intptr_t is, const uint16_t *const *ra, const uint16_t *const *rb, int nref,
                     intptr_t rs, uint32_t w, uint32_t h, bool satd, uint32_t *out, const char *fn)
{
    for (int k = 0; k < nref; k++) out[k] = UINT32_MAX;
    if (!check_wh(w, h, 128, 128, fn)) return;
    if (satd && ((w & 1) || (h & 1))) {
        set_error(XG_ERR_ARG, "%s: SATD needs even sizes (%ux%u)", fn, w, h);
        return;
    }
    const size_t npix = (size_t)w * h, blk = align_up(npix * 2, 256);
    const size_t bytes = blk * (1 + 2 * nref) + 256;
    CallCtx *c = call_ctx(bytes, bytes);
    if (!c) return;
    // input, then the A blocks, then the B blocks
    pack2d(c->host, in, is, (int)w, (int)h, 2);
    for (int k = 0; k < nref; k++) {
        pack2d(c->host + blk * (1 + k), ra[k], rs, (int)w, (int)h, 2);
        pack2d(c->host + blk * (1 + nref + k), rb[k], rs, (int)w, (int)h, 2);
    }
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, blk * (1 + 2 * nref), hipMemcpyHostToDevice, c->stream))) return;
    uint16_t *dA = (uint16_t *)(c->dev + blk), *dB = (uint16_t *)(c->dev + blk * (1 + nref));
    k_avg2<<<dim3((unsigned)std::min<size_t>((npix + 255) / 256, 64), nref), 256, 0, c->stream>>>(
        dA, dB, (int)npix, (int)(blk / 2));
    uint32_t *dres = (uint32_t *)(c->dev + blk * (1 + 2 * nref));
    if (satd)
        k_satd<<<1, 256, 0, c->stream>>>((uint16_t *)c->dev, dA, (int)w, (int)h, dres);
    else
        k_sad<<<nref, 256, 0, c->stream>>>((uint16_t *)c->dev, dA, (int)npix, (int)(blk / 2), dres);
    if (!XG_HIP(hipGetLastError())) return;
    uint32_t *hres = (uint32_t *)(c->host + blk * (1 + 2 * nref));
    if (!XG_HIP(hipMemcpyAsync(hres, dres, 4 * nref, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    for (int k = 0; k < nref; k++) out[k] = hres[k];
}

static void satd_call(const uint16_t *in, intptr_t is, const uint16_t *pr, intptr_t ps, uint32_t w,
                      uint32_t h, uint32_t *out, const char *fn)
{
    *out = UINT32_MAX;
    if (!check_wh(w, h, 128, 128, fn)) return;
    if ((w & 1) || (h & 1)) {
        set_error(XG_ERR_ARG, "%s: SATD needs even sizes (%ux%u)", fn, w, h);
        return;
    }
    const size_t npix = (size_t)w * h, blk = align_up(npix * 2, 256);
    CallCtx *c = call_ctx(2 * blk + 256, 2 * blk + 256);
    if (!c) return;
    pack2d(c->host, in, is, (int)w, (int)h, 2);
    pack2d(c->host + blk, pr, ps, (int)w, (int)h, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, 2 * blk, hipMemcpyHostToDevice, c->stream))) return;
    k_satd<<<1, 256, 0, c->stream>>>((uint16_t *)c->dev, (uint16_t *)(c->dev + blk), (int)w, (int)h,
                                     (uint32_t *)(c->dev + 2 * blk));
    if (!XG_HIP(hipGetLastError())) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + 2 * blk, c->dev + 2 * blk, 4, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    *out = *(uint32_t *)(c->host + 2 * blk);
}

// Device-side 1-D pass: `in` packed [line][N] -> `out` packed [N][line].
static bool fdct1d_dev(CallCtx *c, const int32_t *din, int type, int N, int line, int shift, int32_t *dout)
{
    dim3 grid((line + 63) / 64, N), block(64);
    k_fdct1d<<<grid, block, 0, c->stream>>>(din, type, N, line, shift, dout);
    return XG_HIP(hipGetLastError());
}

static void fdct1d_call(int type, int N, const int32_t *in, intptr_t is, int32_t *out, intptr_t os,
                        int32_t line, int32_t shift)
{
    if (line <= 0 || line > 128 || shift < 0 || shift > 30) {
        set_error(XG_ERR_ARG, "FDct 1-D: bad line %d or shift %d", line, shift);
        return;
    }
    const size_t inB = align_up((size_t)line * N * 4, 256), outB = (size_t)line * N * 4;
    CallCtx *c = call_ctx(inB + outB, inB + outB);
    if (!c) return;
    pack2d(c->host, in, is, N, line, 4);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, (size_t)line * N * 4, hipMemcpyHostToDevice, c->stream))) return;
    if (!fdct1d_dev(c, (int32_t *)c->dev, type, N, line, shift, (int32_t *)(c->dev + inB))) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + inB, c->dev + inB, outB, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    unpack2d(out, os, c->host + inB, line, N, 4);
}

static int lg2(uint32_t v)
{
    int l = 0;
    while ((1u << l) < v) l++;
    return l;
}

// Two-pass transform on device with the layout of h266_forward_trans.c:114-128.
static int32_t fdct2d_call(int hor, int ver, const int32_t *in, intptr_t is, int32_t *out, intptr_t os,
                           uint32_t w, uint32_t h, int32_t s1, int32_t s2)
{
    const size_t n = (size_t)w * h, blk = align_up(n * 4, 256);
    CallCtx *c = call_ctx(3 * blk, 2 * blk);
    if (!c) return XG_ERR_HIP;
    pack2d(c->host, in, is, (int)w, (int)h, 4);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, n * 4, hipMemcpyHostToDevice, c->stream))) return XG_ERR_HIP;
    int32_t *d0 = (int32_t *)c->dev, *d1 = (int32_t *)(c->dev + blk), *d2 = (int32_t *)(c->dev + 2 * blk);
    if (!fdct1d_dev(c, d0, hor, (int)w, (int)h, s1, d1)) return XG_ERR_HIP;   // tmp[k*h + y]
    if (!fdct1d_dev(c, d1, ver, (int)h, (int)w, s2, d2)) return XG_ERR_HIP;   // out[j*w + k]
    if (!XG_HIP(hipMemcpyAsync(c->host + blk, d2, n * 4, hipMemcpyDeviceToHost, c->stream))) return XG_ERR_HIP;
    if (!finish(c)) return XG_ERR_HIP;
    unpack2d(out, os, c->host + blk, (int)w, (int)h, 4);
    return XG_OK;
}

// Packs the rows x cols corner of a strided source into a dstW-wide packed
// buffer: the inverse passes read nothing else (coefficients past the
// zero-out and skipped lines), so nothing else is read from the caller.
static void pack_corner(char *dst, const int32_t *src, intptr_t stride, int rows, int cols, int dstW)
{
    for (int r = 0; r < rows; r++) memcpy(dst + (size_t)r * dstW * 4, src + (intptr_t)r * stride, (size_t)cols * 4);
}

// Device-side inverse pass: `in` packed [N][line] -> `out` packed [line][N].
static bool idct1d_dev(CallCtx *c, const int32_t *din, int type, int N, int line, int skip, int shift,
                       int32_t *dout)
{
    dim3 grid((line + 63) / 64, N), block(64);
    k_idct1d<<<grid, block, 0, c->stream>>>(din, type, N, line, skip, shift, dout);
    return XG_HIP(hipGetLastError());
}

static void idct1d_call(int type, int N, const int32_t *in, intptr_t is, int32_t *out, intptr_t os,
                        int32_t line, int32_t skip, int32_t shift)
{
    if (line <= 0 || line > 128 || skip < 0 || skip > line || shift < 0 || shift > 30) {
        set_error(XG_ERR_ARG, "IDct 1-D: bad line %d, skipLine %d or shift %d", line, skip, shift);
        return;
    }
    const size_t inB = align_up((size_t)line * N * 4, 256), outB = (size_t)line * N * 4;
    CallCtx *c = call_ctx(inB + outB, inB + outB);
    if (!c) return;
    pack_corner(c->host, in, is, tr_kept(type, N), line - skip, line);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, (size_t)line * N * 4, hipMemcpyHostToDevice, c->stream))) return;
    if (!idct1d_dev(c, (int32_t *)c->dev, type, N, line, skip, shift, (int32_t *)(c->dev + inB))) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + inB, c->dev + inB, outB, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    unpack2d(out, os, c->host + inB, N, line, 4);
}

// Inverse of one TU on device: coefficients d0 packed [h][w] -> residual
// packed [h][w] in d2 (d1 holds the transposed first pass).  hor == kSKIP
// runs the transform-skip shift instead.
static bool itx_dev(CallCtx *c, int hor, int ver, uint32_t w, uint32_t h, int skipW, int s1, int s2,
                    const int32_t *d0, int32_t *d1, int32_t *d2)
{
    if (hor == kSKIP) {
        const int n = (int)(w * h), shift = 15 - 8 - ((lg2(w) + lg2(h)) >> 1);
        k_iskip<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(d0, n, shift, d2);
        return XG_HIP(hipGetLastError());
    }
    return idct1d_dev(c, d0, ver, (int)h, (int)w, skipW, s1, d1) &&   // tmp[x*h + y]
           idct1d_dev(c, d1, hor, (int)w, (int)h, 0, s2, d2);          // res[y*w + x]
}

// Rows and columns of a TU's coefficients that the inverse reads.
static void itx_read_region(int hor, int ver, uint32_t w, uint32_t h, int skipW, int *rows, int *cols)
{
    if (hor == kSKIP) {
        *rows = (int)h;
        *cols = (int)w;
        return;
    }
    *rows = tr_kept(ver, (int)h);
    *cols = (int)w - skipW;
}

static int32_t itx_call(int hor, int ver, const int32_t *in, intptr_t is, int32_t *out, intptr_t os, uint32_t w,
                        uint32_t h, int skipW, int32_t s1, int32_t s2)
{
    const size_t n = (size_t)w * h, blk = align_up(n * 4, 256);
    CallCtx *c = call_ctx(3 * blk, 2 * blk);
    if (!c) return XG_ERR_HIP;
    int rows, cols;
    itx_read_region(hor, ver, w, h, skipW, &rows, &cols);
    pack_corner(c->host, in, is, rows, cols, (int)w);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, n * 4, hipMemcpyHostToDevice, c->stream))) return XG_ERR_HIP;
    int32_t *d0 = (int32_t *)c->dev, *d1 = (int32_t *)(c->dev + blk), *d2 = (int32_t *)(c->dev + 2 * blk);
    if (!itx_dev(c, hor, ver, w, h, skipW, s1, s2, d0, d1, d2)) return XG_ERR_HIP;
    if (!XG_HIP(hipMemcpyAsync(c->host + blk, d2, n * 4, hipMemcpyDeviceToHost, c->stream))) return XG_ERR_HIP;
    if (!finish(c)) return XG_ERR_HIP;
    unpack2d(out, os, c->host + blk, (int)w, (int)h, 4);
    return XG_OK;
}

static bool pow2_in(uint32_t v, uint32_t lo, uint32_t hi) { return v >= lo && v <= hi && !(v & (v - 1)); }

// Transform types, zero-out skip and shifts of Xin266IDctWxH
// (h266_encoder/source/h266_inverse_trans.c:115-199) for 10-bit samples.
struct ItxPlan {
    int hor, ver, skipW, s1, s2;
};

static bool itx_plan(uint32_t w, uint32_t h, uint32_t mtsIdx, int32_t isIntraLuma, int32_t enableMts,
                     const char *fn, ItxPlan *p)
{
    static const int8_t tab[6][2] = {{kDCT2, kDCT2}, {kSKIP, kSKIP}, {kDST7, kDST7},
                                     {kDCT8, kDST7}, {kDST7, kDCT8}, {kDCT8, kDCT8}};
    if (mtsIdx > 5 || !pow2_in(w, 1, 128) || !pow2_in(h, 1, 128)) {
        set_error(XG_ERR_ARG, "%s: unsupported %ux%u mts %u", fn, w, h, mtsIdx);
        return false;
    }
    p->hor = tab[mtsIdx][0];
    p->ver = tab[mtsIdx][1];
    p->s1 = 6 + 1;                                                   // :145
    p->s2 = 6 + 15 - 1 - 10;                                         // :146
    p->skipW = (p->hor != kDCT2 && w == 32) ? 16 : (w > 32 ? (int)w - 32 : 0);   // :147
    if (isIntraLuma && enableMts && mtsIdx == 0) {                    // :149-159
        if (w >= 4 && w <= 16) p->hor = kDST7;
        if (h >= 4 && h <= 16) p->ver = kDST7;
    }
    if (p->hor == kSKIP) return true;                                 // :173-182, any size to 128
    if (w < 2 || h < 2 || w > 64 || h > 64) {
        // One-sided blocks (:201-228) leave their result in the reference's
        // section scratch buffer, so there is no output to reproduce.
        set_error(XG_ERR_ARG, "%s: unsupported %ux%u for a transform", fn, w, h);
        return false;
    }
    if ((p->hor != kDCT2 && (w < 4 || w > 32)) || (p->ver != kDCT2 && (h < 4 || h > 32))) {
        set_error(XG_ERR_ARG, "%s: DST-VII/DCT-VIII need sizes 4..32", fn);
        return false;
    }
    return true;
}

}  // namespace xg

using namespace xg;

extern "C" {

// ---- distortion ----
void Xg266ComputeSad(xg_pixel *input, intptr_t inputStride, xg_pixel *pred, intptr_t predStride,
                     uint32_t width, uint32_t height, uint32_t *sad)
{
    const uint16_t *r[1] = {pred};
    sad_multi(input, inputStride, r, 1, predStride, width, height, sad, "ComputeSad");
}

void Xg266ComputeSatd(xg_pixel *input, intptr_t inputStride, xg_pixel *pred, intptr_t predStride,
                      uint32_t width, uint32_t height, uint32_t *satd)
{
    satd_call(input, inputStride, pred, predStride, width, height, satd, "ComputeSatd");
}

void Xg266ComputeSadx8(xg_pixel *input, intptr_t inputStride, xg_pixel *ref[8], intptr_t refStride,
                       uint32_t width, uint32_t height, uint32_t *sad)
{
    sad_multi(input, inputStride, (const uint16_t *const *)ref, 8, refStride, width, height, sad, "ComputeSadx8");
}

void Xg266ComputeSadx5(xg_pixel *input, intptr_t inputStride, xg_pixel *ref0, xg_pixel *ref1,
                       xg_pixel *ref2, xg_pixel *ref3, xg_pixel *ref4, intptr_t refStride,
                       uint32_t width, uint32_t height, uint32_t *sad)
{
    const uint16_t *r[5] = {ref0, ref1, ref2, ref3, ref4};
    sad_multi(input, inputStride, r, 5, refStride, width, height, sad, "ComputeSadx5");
}

void Xg266ComputeSadx3(xg_pixel *input, intptr_t inputStride, xg_pixel *ref0, xg_pixel *ref1,
                       xg_pixel *ref2, intptr_t refStride, uint32_t width, uint32_t height, uint32_t *sad)
{
    const uint16_t *r[3] = {ref0, ref1, ref2};
    sad_multi(input, inputStride, r, 3, refStride, width, height, sad, "ComputeSadx3");
}

void Xg266ComputeAvgSad(xg_pixel *input, intptr_t inputStride, xg_pixel *refA, xg_pixel *refB,
                        intptr_t refStride, uint32_t width, uint32_t height, uint32_t *sad)
{
    const uint16_t *a[1] = {refA}, *b[1] = {refB};
    avg_dist(input, inputStride, a, b, 1, refStride, width, height, false, sad, "ComputeAvgSad");
}

void Xg266ComputeAvgSatd(xg_pixel *input, intptr_t inputStride, xg_pixel *refA, xg_pixel *refB,
                         intptr_t refStride, uint32_t width, uint32_t height, uint32_t *satd)
{
    const uint16_t *a[1] = {refA}, *b[1] = {refB};
    avg_dist(input, inputStride, a, b, 1, refStride, width, height, true, satd, "ComputeAvgSatd");
}

void Xg266ComputeAvgSadx8(xg_pixel *input, intptr_t inputStride, xg_pixel *refA[8], xg_pixel *refB[8],
                          intptr_t refStride, uint32_t width, uint32_t height, uint32_t *sad)
{
    avg_dist(input, inputStride, (const uint16_t *const *)refA, (const uint16_t *const *)refB, 8, refStride, width,
             height, false, sad, "ComputeAvgSadx8");
}

void Xg266QuantInvQuant(xg_coeff *qCoeff, xg_coeff *tCoeff, xg_coeff *rCoeff, intptr_t coeffStride, uint32_t width,
                        uint32_t height, uint32_t cgWidth, uint32_t cgHeight, int32_t qMult, int32_t qAdd,
                        int32_t qShift, int32_t iqMult, int32_t iqAdd, int32_t iqShift, uint64_t *nzCGBitMapRs)
{
    const char *fn = "QuantInvQuant";
    if (!qCoeff || !tCoeff || !rCoeff || !nzCGBitMapRs || !check_wh(width, height, 128, 128, fn)) return;
    if (cgWidth == 0 || cgHeight == 0 || width % cgWidth || height % cgHeight ||
        (width / cgWidth) * (height / cgHeight) > 64 || qShift < 0 || qShift > 31 || iqShift < 0 || iqShift > 31) {
        set_error(XG_ERR_ARG, "%s: %ux%u in %ux%u groups (at most 64) or shifts %d/%d out of range", fn, width, height,
                  cgWidth, cgHeight, qShift, iqShift);
        return;
    }
    const size_t n = (size_t)width * height, blk = align_up(n * 4, 256);
    CallCtx *c = call_ctx(3 * blk + 256, 3 * blk + 256);
    if (!c) return;
    pack2d(c->host, tCoeff, coeffStride, (int)width, (int)height, 4);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, n * 4, hipMemcpyHostToDevice, c->stream))) return;
    int32_t *dq = (int32_t *)(c->dev + blk), *dr = (int32_t *)(c->dev + 2 * blk);
    unsigned long long *dm = (unsigned long long *)(c->dev + 3 * blk);
    k_qiq<<<1, 64, 0, c->stream>>>((const int32_t *)c->dev, dq, dr, (int)width, (int)height, (int)cgWidth,
                                   (int)cgHeight, qMult, qAdd, qShift, iqMult, iqAdd, iqShift, dm);
    if (!XG_HIP(hipGetLastError())) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + blk, c->dev + blk, 2 * blk + 8, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    unpack2d(qCoeff, coeffStride, c->host + blk, (int)width, (int)height, 4);
    unpack2d(rCoeff, coeffStride, c->host + 2 * blk, (int)width, (int)height, 4);
    *nzCGBitMapRs = *(const uint64_t *)(c->host + 3 * blk);
}

void Xg266ComputeHada8x8(xg_pixel *input, intptr_t inputStride, uint32_t *satd)
{
    static const xg_pixel zero[8] = {0};
    satd_call(input, inputStride, (xg_pixel *)zero, 0, 8, 8, satd, "ComputeHada8x8");
}

// ---- luma interpolation ----
#define XG_LUMA_U16(NAME, DIR, FX, FY)                                                              \
    void NAME(const xg_pixel *src, intptr_t ss, xg_pixel *dst, intptr_t ds, int32_t frac,          \
              uint32_t w, uint32_t h, uint32_t fi)                                                  \
    {                                                                                               \
        interp_u16(8, false, DIR, src, ss, dst, ds, FX, FY, (int)fi, w, h, #NAME);                  \
    }
#define XG_LUMA_S16(NAME, DIR, FX, FY)                                                              \
    void NAME(const xg_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,           \
              uint32_t w, uint32_t h, uint32_t fi)                                                  \
    {                                                                                               \
        interp_u16(8, true, DIR, src, ss, dst, ds, FX, FY, (int)fi, w, h, #NAME);                   \
    }
XG_LUMA_U16(Xg266InterpCopy, kCopy, 0, 0)
XG_LUMA_U16(Xg266LumaInterpHor, kHor, frac & 15, 0)
XG_LUMA_U16(Xg266LumaInterpVet, kVet, 0, (frac >> 4) & 15)
XG_LUMA_U16(Xg266LumaInterpHorVet, kHorVet, frac & 15, (frac >> 4) & 15)
XG_LUMA_S16(Xg266InterpCopyU8S16, kCopy, 0, 0)
XG_LUMA_S16(Xg266LumaInterpHorU8S16, kHor, frac & 15, 0)
XG_LUMA_S16(Xg266LumaInterpVetU8S16, kVet, 0, (frac >> 4) & 15)
XG_LUMA_S16(Xg266LumaInterpHorVetU8S16, kHorVet, frac & 15, (frac >> 4) & 15)

void Xg266LumaInterpVetS16U8(const int16_t *src, intptr_t ss, xg_pixel *dst, intptr_t ds, int32_t frac,
                             uint32_t w, uint32_t h, uint32_t fi)
{
    interp_s16(8, false, src, ss, dst, ds, (frac >> 4) & 15, (int)fi, w, h, "Xg266LumaInterpVetS16U8");
}
void Xg266LumaInterpVetS16S16(const int16_t *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                              uint32_t w, uint32_t h, uint32_t fi)
{
    interp_s16(8, true, src, ss, dst, ds, (frac >> 4) & 15, (int)fi, w, h, "Xg266LumaInterpVetS16S16");
}

// ---- chroma interpolation (frac = (fy << 5) | fx) ----
#define XG_CHROMA_U16(NAME, DIR, FX, FY)                                                            \
    void NAME(const xg_pixel *src, intptr_t ss, xg_pixel *dst, intptr_t ds, int32_t frac,          \
              uint32_t w, uint32_t h, uint32_t fi)                                                  \
    {                                                                                               \
        (void)fi;                                                                                   \
        interp_u16(4, false, DIR, src, ss, dst, ds, FX, FY, 0, w, h, #NAME);                        \
    }
#define XG_CHROMA_S16(NAME, DIR, FX, FY)                                                            \
    void NAME(const xg_pixel *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,           \
              uint32_t w, uint32_t h, uint32_t fi)                                                  \
    {                                                                                               \
        (void)fi;                                                                                   \
        interp_u16(4, true, DIR, src, ss, dst, ds, FX, FY, 0, w, h, #NAME);                         \
    }
XG_CHROMA_U16(Xg266ChromaInterpHor, kHor, frac & 31, 0)
XG_CHROMA_U16(Xg266ChromaInterpVet, kVet, 0, (frac >> 5) & 31)
XG_CHROMA_U16(Xg266ChromaInterpHorVet, kHorVet, frac & 31, (frac >> 5) & 31)
XG_CHROMA_S16(Xg266ChromaInterpHorU8S16, kHor, frac & 31, 0)
XG_CHROMA_S16(Xg266ChromaInterpVetU8S16, kVet, 0, (frac >> 5) & 31)
XG_CHROMA_S16(Xg266ChromaInterpHorVetU8S16, kHorVet, frac & 31, (frac >> 5) & 31)

void Xg266ChromaInterpVetS16U8(const int16_t *src, intptr_t ss, xg_pixel *dst, intptr_t ds, int32_t frac,
                               uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    interp_s16(4, false, src, ss, dst, ds, (frac >> 5) & 31, 0, w, h, "Xg266ChromaInterpVetS16U8");
}
void Xg266ChromaInterpVetS16S16(const int16_t *src, intptr_t ss, int16_t *dst, intptr_t ds, int32_t frac,
                                uint32_t w, uint32_t h, uint32_t fi)
{
    (void)fi;
    interp_s16(4, true, src, ss, dst, ds, (frac >> 5) & 31, 0, w, h, "Xg266ChromaInterpVetS16S16");
}

void Xg266InterpAvgS16U8(const int16_t *src0, intptr_t src0Stride, const int16_t *src1, intptr_t src1Stride,
                         xg_pixel *dst, intptr_t dstStride, uint32_t width, uint32_t height)
{
    bipred(src0, src0Stride, src1, src1Stride, dst, dstStride, width, height, 0, 0, 0, "Xg266InterpAvgS16U8");
}

void Xg266InterpWeightS16U8(const int16_t *src0, intptr_t src0Stride, const int16_t *src1, intptr_t src1Stride,
                            xg_pixel *dst, intptr_t dstStride, uint32_t width, uint32_t height,
                            int32_t weightA, int32_t weightB)
{
    bipred(src0, src0Stride, src1, src1Stride, dst, dstStride, width, height, 1, weightA, weightB,
           "Xg266InterpWeightS16U8");
}

// ---- forward transforms ----
#define XG_FDCT(NAME, TYPE, N)                                                                      \
    void NAME(xg_coeff *in, intptr_t is, xg_coeff *out, intptr_t os, int32_t line, int32_t shift)  \
    {                                                                                               \
        fdct1d_call(TYPE, N, in, is, out, os, line, shift);                                         \
    }
XG_FDCT(Xg26xFDct2P2, kDCT2, 2)
XG_FDCT(Xg26xFDct2P4, kDCT2, 4)
XG_FDCT(Xg26xFDct2P8, kDCT2, 8)
XG_FDCT(Xg26xFDct2P16, kDCT2, 16)
XG_FDCT(Xg26xFDct2P32, kDCT2, 32)
XG_FDCT(Xg26xFDct2P64, kDCT2, 64)
XG_FDCT(Xg26xFDst7P4, kDST7, 4)
XG_FDCT(Xg26xFDst7P8, kDST7, 8)
XG_FDCT(Xg26xFDst7P16, kDST7, 16)
XG_FDCT(Xg26xFDst7P32, kDST7, 32)
XG_FDCT(Xg26xFDct8P4, kDCT8, 4)
XG_FDCT(Xg26xFDct8P8, kDCT8, 8)
XG_FDCT(Xg26xFDct8P16, kDCT8, 16)
XG_FDCT(Xg26xFDct8P32, kDCT8, 32)

void Xg26xFDct2WxH(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                   uint32_t width, uint32_t height, xg_coeff *tempBuf, int32_t shift1st, int32_t shift2nd)
{
    (void)tempBuf;
    if (width < 2 || height < 2 || width > 64 || height > 64 || (width & (width - 1)) || (height & (height - 1))) {
        set_error(XG_ERR_ARG, "Xg26xFDct2WxH: unsupported %ux%u", width, height);
        return;
    }
    fdct2d_call(kDCT2, kDCT2, input, inputStride, output, outputStride, width, height, shift1st, shift2nd);
}

void Xg26xBlockSubForDct(xg_pixel *input, intptr_t inputStride, xg_pixel *recon, intptr_t reconStride,
                         xg_coeff *diff, intptr_t diffStride, uint32_t width, uint32_t height)
{
    if (!check_wh(width, height, 128, 128, "Xg26xBlockSubForDct")) return;
    const size_t n = (size_t)width * height, blk = align_up(n * 4, 256);
    CallCtx *c = call_ctx(3 * blk, 3 * blk);
    if (!c) return;
    pack2d(c->host, input, inputStride, (int)width, (int)height, 2);
    pack2d(c->host + blk, recon, reconStride, (int)width, (int)height, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, 2 * blk, hipMemcpyHostToDevice, c->stream))) return;
    k_blocksub<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>((uint16_t *)c->dev, (uint16_t *)(c->dev + blk),
                                                                 (int)n, (int32_t *)(c->dev + 2 * blk));
    if (!XG_HIP(hipGetLastError())) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + 2 * blk, c->dev + 2 * blk, n * 4, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    unpack2d(diff, diffStride, c->host + 2 * blk, (int)width, (int)height, 4);
}

int32_t Xg266FDctWxH(const xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                     uint32_t width, uint32_t height, uint32_t mtsIdx, int32_t isIntraLuma, int32_t enableMts)
{
    static const int8_t tab[6][2] = {{kDCT2, kDCT2}, {kSKIP, kSKIP}, {kDST7, kDST7},
                                     {kDCT8, kDST7}, {kDST7, kDCT8}, {kDCT8, kDCT8}};
    if (width < 2 || height < 2 || width > 64 || height > 64 || (width & (width - 1)) ||
        (height & (height - 1)) || mtsIdx > 5 || mtsIdx == 1) {
        set_error(XG_ERR_ARG, "Xg266FDctWxH: unsupported %ux%u mts %u", width, height, mtsIdx);
        return XG_ERR_ARG;
    }
    int hor = tab[mtsIdx][0], ver = tab[mtsIdx][1];
    if (isIntraLuma && enableMts && mtsIdx == 0) {         // h266_forward_trans.c:75-86
        if (width >= 4 && width <= 16) hor = kDST7;
        if (height >= 4 && height <= 16) ver = kDST7;
    }
    if ((hor != kDCT2 && (width < 4 || width > 32)) || (ver != kDCT2 && (height < 4 || height > 32))) {
        set_error(XG_ERR_ARG, "Xg266FDctWxH: DST-VII/DCT-VIII need sizes 4..32");
        return XG_ERR_ARG;
    }
    const int lw = lg2(width), lh = lg2(height);
    return fdct2d_call(hor, ver, input, inputStride, output, outputStride, width, height,
                       lw + 10 + 6 - 15, lh + 6);                         // :72-73
}

// ---- inverse transforms and reconstruction ----
#define XG_IDCT(NAME, TYPE, N)                                                                          \
    void NAME(xg_coeff *in, intptr_t is, xg_coeff *out, intptr_t os, int32_t line, int32_t skipLine,   \
              int32_t shift)                                                                            \
    {                                                                                                   \
        idct1d_call(TYPE, N, in, is, out, os, line, skipLine, shift);                                   \
    }
XG_IDCT(Xg26xIDct2P2, kDCT2, 2)
XG_IDCT(Xg26xIDct2P4, kDCT2, 4)
XG_IDCT(Xg26xIDct2P8, kDCT2, 8)
XG_IDCT(Xg26xIDct2P16, kDCT2, 16)
XG_IDCT(Xg26xIDct2P32, kDCT2, 32)
XG_IDCT(Xg26xIDct2P64, kDCT2, 64)
XG_IDCT(Xg26xIDst7P4, kDST7, 4)
XG_IDCT(Xg26xIDst7P8, kDST7, 8)
XG_IDCT(Xg26xIDst7P16, kDST7, 16)
XG_IDCT(Xg26xIDst7P32, kDST7, 32)
XG_IDCT(Xg26xIDct8P4, kDCT8, 4)
XG_IDCT(Xg26xIDct8P8, kDCT8, 8)
XG_IDCT(Xg26xIDct8P16, kDCT8, 16)
XG_IDCT(Xg26xIDct8P32, kDCT8, 32)

void Xg266IDct2WxH(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                   uint32_t width, uint32_t height, xg_coeff *tempBuf, int32_t shift1st, int32_t shift2nd)
{
    (void)tempBuf;
    if (!pow2_in(width, 2, 64) || !pow2_in(height, 2, 64) || shift1st < 0 || shift1st > 30 || shift2nd < 0 ||
        shift2nd > 30) {
        set_error(XG_ERR_ARG, "Xg266IDct2WxH: unsupported %ux%u or shifts %d/%d", width, height, shift1st,
                  shift2nd);
        return;
    }
    itx_call(kDCT2, kDCT2, input, inputStride, output, outputStride, width, height,
             width > 32 ? (int)width - 32 : 0, shift1st, shift2nd);
}

void Xg266ISkipWxH(xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                   uint32_t width, uint32_t height)
{
    if (!pow2_in(width, 1, 128) || !pow2_in(height, 1, 128)) {
        set_error(XG_ERR_ARG, "Xg266ISkipWxH: unsupported %ux%u", width, height);
        return;
    }
    itx_call(kSKIP, kSKIP, input, inputStride, output, outputStride, width, height, 0, 0, 0);
}

int32_t Xg266IDctWxH(const xg_coeff *input, intptr_t inputStride, xg_coeff *output, intptr_t outputStride,
                     uint32_t width, uint32_t height, uint32_t mtsIdx, int32_t isIntraLuma, int32_t enableMts)
{
    ItxPlan p;
    if (!itx_plan(width, height, mtsIdx, isIntraLuma, enableMts, "Xg266IDctWxH", &p)) return XG_ERR_ARG;
    return itx_call(p.hor, p.ver, input, inputStride, output, outputStride, width, height, p.skipW, p.s1, p.s2);
}

void Xg26xBlockRecon(xg_coeff *input, intptr_t inputStride, xg_pixel *pred, intptr_t predStride, xg_pixel *recon,
                     intptr_t reconStride, uint32_t width, uint32_t height)
{
    if (!check_wh(width, height, 128, 128, "Xg26xBlockRecon")) return;
    const size_t n = (size_t)width * height, blk = align_up(n * 4, 256);
    CallCtx *c = call_ctx(3 * blk, 3 * blk);
    if (!c) return;
    pack2d(c->host, input, inputStride, (int)width, (int)height, 4);
    pack2d(c->host + blk, pred, predStride, (int)width, (int)height, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, 2 * blk, hipMemcpyHostToDevice, c->stream))) return;
    k_recon<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>((int32_t *)c->dev, (uint16_t *)(c->dev + blk),
                                                              (int)n, (uint16_t *)(c->dev + 2 * blk));
    if (!XG_HIP(hipGetLastError())) return;
    if (!XG_HIP(hipMemcpyAsync(c->host + 2 * blk, c->dev + 2 * blk, n * 2, hipMemcpyDeviceToHost, c->stream))) return;
    if (!finish(c)) return;
    unpack2d(recon, reconStride, c->host + 2 * blk, (int)width, (int)height, 2);
}

int32_t Xg266IDctAdd(const xg_coeff *input, intptr_t inputStride, const xg_pixel *pred, intptr_t predStride,
                     xg_pixel *output, intptr_t outputStride, uint32_t width, uint32_t height, uint32_t mtsIdx,
                     int32_t isIntraLuma, int32_t enableMts)
{
    ItxPlan p;
    if (!itx_plan(width, height, mtsIdx, isIntraLuma, enableMts, "Xg266IDctAdd", &p)) return XG_ERR_ARG;
    if (width > 64 || height > 64) {                    // the reference's 64x64 residual buffer
        set_error(XG_ERR_ARG, "Xg266IDctAdd: unsupported %ux%u", width, height);
        return XG_ERR_ARG;
    }
    const size_t n = (size_t)width * height, blk = align_up(n * 4, 256);
    CallCtx *c = call_ctx(5 * blk, 3 * blk);
    if (!c) return XG_ERR_HIP;
    int rows, cols;
    itx_read_region(p.hor, p.ver, width, height, p.skipW, &rows, &cols);
    pack_corner(c->host, input, inputStride, rows, cols, (int)width);
    pack2d(c->host + blk, pred, predStride, (int)width, (int)height, 2);
    if (!XG_HIP(hipMemcpyAsync(c->dev, c->host, 2 * blk, hipMemcpyHostToDevice, c->stream))) return XG_ERR_HIP;
    int32_t *d0 = (int32_t *)c->dev, *d1 = (int32_t *)(c->dev + 2 * blk), *d2 = (int32_t *)(c->dev + 3 * blk);
    uint16_t *dp = (uint16_t *)(c->dev + blk), *dr = (uint16_t *)(c->dev + 4 * blk);
    if (!itx_dev(c, p.hor, p.ver, width, height, p.skipW, p.s1, p.s2, d0, d1, d2)) return XG_ERR_HIP;
    k_recon<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(d2, dp, (int)n, dr);
    if (!XG_HIP(hipGetLastError())) return XG_ERR_HIP;
    if (!XG_HIP(hipMemcpyAsync(c->host + 2 * blk, dr, n * 2, hipMemcpyDeviceToHost, c->stream))) return XG_ERR_HIP;
    if (!finish(c)) return XG_ERR_HIP;
    unpack2d(output, outputStride, c->host + 2 * blk, (int)width, (int)height, 2);
    return XG_OK;
}

}  // extern "C"

extern "C" {

// Table population following h266_encoder/source/h266_func_init.c:87-690 (C
// defaults: SATD NULL for widths 1 and 2 at :648-649; DST-VII/DCT-VIII 4..32
// and DCT-II 2..64 at :349-374; 2-D DCT-II for every [lgW][lgH] at :72-85)
// followed by the SAD/SATD binding of Xin266FuncUpdate
// (h266_encoder/source/h266_encoder_create.c:3332-3400).
int32_t Xg266FuncInit(xg_func_struct *f, int32_t satdMd)
{
    if (!f) {
        set_error(XG_ERR_ARG, "Xg266FuncInit: NULL table");
        return XG_ERR_ARG;
    }
    memset(f, 0, sizeof(*f));
    if (!device_init()) return XG_ERR_HIP;
    static const XgFor1dDct dct2[XG_BLOCK_NUM] = {nullptr, Xg26xFDct2P2, Xg26xFDct2P4, Xg26xFDct2P8,
                                                  Xg26xFDct2P16, Xg26xFDct2P32, Xg26xFDct2P64, nullptr};
    static const XgFor1dDct dct8[XG_BLOCK_NUM] = {nullptr, nullptr, Xg26xFDct8P4, Xg26xFDct8P8,
                                                  Xg26xFDct8P16, Xg26xFDct8P32, nullptr, nullptr};
    static const XgFor1dDct dst7[XG_BLOCK_NUM] = {nullptr, nullptr, Xg26xFDst7P4, Xg26xFDst7P8,
                                                  Xg26xFDst7P16, Xg26xFDst7P32, nullptr, nullptr};
    static const XgInv1dDct idct2[XG_BLOCK_NUM] = {nullptr, Xg26xIDct2P2, Xg26xIDct2P4, Xg26xIDct2P8,
                                                   Xg26xIDct2P16, Xg26xIDct2P32, Xg26xIDct2P64, nullptr};
    static const XgInv1dDct idct8[XG_BLOCK_NUM] = {nullptr, nullptr, Xg26xIDct8P4, Xg26xIDct8P8,
                                                   Xg26xIDct8P16, Xg26xIDct8P32, nullptr, nullptr};
    static const XgInv1dDct idst7[XG_BLOCK_NUM] = {nullptr, nullptr, Xg26xIDst7P4, Xg26xIDst7P8,
                                                   Xg26xIDst7P16, Xg26xIDst7P32, nullptr, nullptr};
    for (int i = 0; i < XG_BLOCK_NUM; i++) {
        f->pfXinComputeSad[i] = Xg266ComputeSad;
        f->pfXinComputeSatd[i] = i >= 2 ? Xg266ComputeSatd : nullptr;
        f->pfXinComputeDist[i] = satdMd ? f->pfXinComputeSatd[i] : f->pfXinComputeSad[i];
        f->pfXinComputeAvgSad[i] = i >= 2 ? Xg266ComputeAvgSad : nullptr;
        f->pfXinComputeAvgSatd[i] = i >= 2 ? Xg266ComputeAvgSatd : nullptr;
        f->pfXinComputeAvgDist[i] = satdMd ? f->pfXinComputeAvgSatd[i] : f->pfXinComputeAvgSad[i];
        f->pfXinComputeSadx8[i] = Xg266ComputeSadx8;
        f->pfXinComputeAvgSadx8[i] = Xg266ComputeAvgSadx8;
        f->pfXinQuantInvQuant[i] = i < 7 ? Xg266QuantInvQuant : nullptr;
        f->pfXinComputeSadx5[i] = Xg266ComputeSadx5;
        f->pfXinComputeSadx3[i] = Xg266ComputeSadx3;
        f->pfXinLumaInterp[0][0][i] = Xg266InterpCopy;
        f->pfXinLumaInterp[1][0][i] = Xg266LumaInterpHor;
        f->pfXinLumaInterp[0][1][i] = Xg266LumaInterpVet;
        f->pfXinLumaInterp[1][1][i] = Xg266LumaInterpHorVet;
        f->pfXinLumaInterpS16[0][0][i] = Xg266InterpCopyU8S16;
        f->pfXinLumaInterpS16[1][0][i] = Xg266LumaInterpHorU8S16;
        f->pfXinLumaInterpS16[0][1][i] = Xg266LumaInterpVetU8S16;
        f->pfXinLumaInterpS16[1][1][i] = Xg266LumaInterpHorVetU8S16;
        f->pfXinChromaInterp[0][0][i] = Xg266InterpCopy;
        f->pfXinChromaInterp[1][0][i] = Xg266ChromaInterpHor;
        f->pfXinChromaInterp[0][1][i] = Xg266ChromaInterpVet;
        f->pfXinChromaInterp[1][1][i] = Xg266ChromaInterpHorVet;
        f->pfXinChromaInterpS16[0][0][i] = Xg266InterpCopyU8S16;
        f->pfXinChromaInterpS16[1][0][i] = Xg266ChromaInterpHorU8S16;
        f->pfXinChromaInterpS16[0][1][i] = Xg266ChromaInterpVetU8S16;
        f->pfXinChromaInterpS16[1][1][i] = Xg266ChromaInterpHorVetU8S16;
        f->pfXinInterpAvg[i] = Xg266InterpAvgS16U8;
        f->pfXinInterpWeight[i] = Xg266InterpWeightS16U8;
        f->pfXinFor1dDct[0][i] = dct2[i];   // XIN_DCT2
        f->pfXinFor1dDct[1][i] = dct8[i];   // XIN_DCT8
        f->pfXinFor1dDct[2][i] = dst7[i];   // XIN_DST7
        f->pfXinBlockSubForDct[i] = Xg26xBlockSubForDct;
        f->pfXinInv1dDct[0][i] = idct2[i];  // h266_func_init.c:376-401
        f->pfXinInv1dDct[1][i] = idct8[i];
        f->pfXinInv1dDct[2][i] = idst7[i];
        f->pfXinInv2dSkip[i] = Xg266ISkipWxH;                              // :410-419
        f->pfXinBlockAddForDct[i] = (i >= 1 && i <= 6) ? Xg26xBlockRecon : nullptr;   // :161-166
        f->pfXinBiliInterp[0][0][i] = Xg266BiliInterpCopy;                  // :496-530
        f->pfXinBiliInterp[1][0][i] = Xg266BiliInterpHor;
        f->pfXinBiliInterp[0][1][i] = Xg266BiliInterpVet;
        f->pfXinBiliInterp[1][1][i] = Xg266BiliInterpHorVet;
        f->pfXinComputeSadS16[i] = Xg26xComputeSadS16;                       // :118-125
        for (int j = 0; j < XG_BLOCK_NUM - 1; j++)
            if (i < XG_BLOCK_NUM - 1) {
                f->pfXinFor2dDct2[i][j] = Xg26xFDct2WxH;
                f->pfXinInv2dDct2[i][j] = Xg266IDct2WxH;                  // :72-85
            }
    }
    f->pfXinComputeHada8x8 = Xg266ComputeHada8x8;
    return XG_OK;
}

}  // extern "C"

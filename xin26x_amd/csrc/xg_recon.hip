// xg_recon.hip -- batched reconstruction: inverse DCT-II / DST-VII / DCT-VIII
// (or transform skip) of dequantised coefficients plus the prediction add, on
// int8 MFMA.
//
// Per block (W x H, int16 coefficients) the inverse of
// h266_encoder/source/h266_inverse_trans.c:115-199 followed by the add of
// h266_idct_add.c:95-103, as H.266 clause 8.7.4 defines the arithmetic:
//   stage 1  E[i][x] = Clip16((sum_{k<nzH} T_H[k][i] * C[k][x] + 64) >> 7),   x < nzW
//   stage 2  R[i][o] = Clip16((sum_{x<nzW} T_W[x][o] * E[i][x] + 512) >> 10)
//   recon    Clip3(0, 1023, pred + R)
// nzW, nzH are the zero-out bounds (32 for DCT-II-64, 16 for 32-point
// DST-VII/DCT-VIII): coefficients past them are not read, so both contractions
// have K <= 32, one MFMA step.  Both data operands (int16) are split into two
// signed int8 limbs with an offset,
//   v = 256*h + l + 128,  h = v >> 8 in [-128,127],  l = (v & 255) - 128 in [-128,127]
// so sum_k T[k][o] v_k = 256 * (T.h) + (T.l) + 128 * sum_{k<nz} T[k][o]: two
// MFMAs, and the last term, a constant per output index, starts in the
// accumulator with the rounding offset.
// |limb partial| <= 128*90*32 = 368,640 and |sum| <= 90*32768*32 < 2^31.
//
// Operand roles: stage 1 has the coefficient columns in A (M = columns (b, x),
// K = k) and the basis in B (N = output row i), so each lane receives four
// consecutive columns x of one row i and packs them straight into the limb
// planes; stage 2 has the basis in A (M = output column o, K = x) and those
// limb rows in B (N = rows (b, i)), so each lane receives four consecutive
// output columns of one row: one 8-byte prediction load and recon store.
// Both bases are read as R[o][k] = T[k][o] (the transposed matrix, zero for
// k >= nz), so padding bytes past nz multiply zeros.
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
#include "xg_runtime.hpp"
#include "xg_trmatrix.hpp"

typedef int v4i __attribute__((ext_vector_type(4)));

namespace xg {

// Transposed, zero-padded int8 bases: for (type, lgN) max(16, N) rows o of 32
// bytes, entry k = T[k][o] for k < nz(type, N), zero elsewhere; then, per
// (type, lgN), max(16, N) int32 limb-offset terms 128 * sum_{k<nz} T[k][o].
struct RecMats { int off[3][7]; int coff[3][7]; int total; };

static std::once_flag g_rmat_once;
static int8_t *g_rmats = nullptr;
static RecMats g_rm;
static bool g_rmat_ok = false;

static bool rec_mats_init()
{
    std::call_once(g_rmat_once, [] {
        int o = 0;
        for (int t = 0; t < 3; t++)
            for (int lg = 1; lg <= 6; lg++) {
                const int N = 1 << lg;
                const bool ok = t == kDCT2 ? true : (N >= 4 && N <= 32);
                g_rm.off[t][lg] = ok ? o : -1;
                if (ok) o += std::max(16, N) * 32;
            }
        for (int t = 0; t < 3; t++)
            for (int lg = 1; lg <= 6; lg++) {
                const int N = 1 << lg;
                g_rm.coff[t][lg] = g_rm.off[t][lg] < 0 ? -1 : o;
                if (g_rm.off[t][lg] >= 0) o += 4 * std::max(16, N);
            }
        g_rm.total = o;
        std::vector<int8_t> h((size_t)o, 0);
        for (int t = 0; t < 3; t++)
            for (int lg = 1; lg <= 6; lg++) {
                if (g_rm.off[t][lg] < 0) continue;
                const int N = 1 << lg, nz = tr_kept(t, N);
                for (int r = 0; r < N; r++) {
                    int32_t c = 0;
                    for (int k = 0; k < nz; k++) {
                        h[(size_t)g_rm.off[t][lg] + (size_t)r * 32 + k] = (int8_t)tr_entry(kTr, t, N, k, r);
                        c += 128 * tr_entry(kTr, t, N, k, r);
                    }
                    memcpy(&h[(size_t)g_rm.coff[t][lg] + 4 * (size_t)r], &c, 4);
                }
            }
        if (!XG_HIP(hipMalloc((void **)&g_rmats, h.size()))) return;
        if (!XG_HIP(hipMemcpy(g_rmats, h.data(), h.size(), hipMemcpyHostToDevice))) return;
        g_rmat_ok = true;
    });
    return g_rmat_ok;
}

constexpr int kRecGroupMax = 64;    // blocks per group: one relative offset per lane
constexpr int kRecStage = 4096;     // int16 coefficients staged per group (rows k < nzH)
constexpr int kRecPlane = 6144;     // bytes per limb plane of the stage-1 output (two planes)
constexpr int kRecSlack = 32;       // stage-2 fragments of the last row reach 24 bytes past it

// limb-plane row stride: 8-byte aligned, long rows padded against bank conflicts
__host__ __device__ inline int rec_stride(int nzW) { return nzW <= 8 ? 8 : nzW + 8; }

struct RecGroup {
    long long base;     // element offset the group's relative block offsets start from
    uint32_t desc;      // lgW | lgH << 4 | horTx << 8 | verTx << 12
    int32_t count;      // blocks, 1..kRecGroupMax, all of this descriptor
    int32_t first;      // index of the group's first relative offset
    int32_t pad;
};

inline int rec_group_cap(int W, int H, int hor, int ver)
{
    if (hor == kSKIP) return kRecGroupMax;
    const int nzW = tr_kept(hor, W), nzH = tr_kept(ver, H);
    int c = kRecStage / (nzH * W);
    c = std::min(c, kRecPlane / (H * rec_stride(nzW)));
    return std::min(c, kRecGroupMax);
}

struct RecArgs {
    const int16_t *coeff;
    const uint16_t *pred;
    uint16_t *recon;
    int n;
    const int8_t *mats;
    int matOff[3][7];
    int constOff[3][7];
};

__device__ __forceinline__ int clip16(int v) { return min(max(v, -32768), 32767); }

// offset limbs of 8 int16 values packed two per dword (value order):
// l = (v & 255) - 128 (the low byte with its top bit flipped), h = v >> 8
__device__ __forceinline__ void limbs2(const uint32_t (&w)[4], long &bl, long &bh)
{
    uint32_t ml[4], mh[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        ml[q] = w[q] ^ 0x00800080u;
        mh[q] = as_u32(as_v2s(w[q]) >> (v2s){8, 8});
    }
    bl = bytes8_lo16(ml);
    bh = bytes8_lo16(mh);
}

constexpr int kRecWaves = 4;                   // waves sharing one group
constexpr int kRecThreads = 64 * kRecWaves;

// Transform skip (h266_inverse_trans.c:91-113 then the add): two samples per
// thread and step, straight from global memory.
__device__ __forceinline__ void recon_skip(const RecArgs &a, const RecGroup &G, int myRel, int lgW, int lgH,
                                           int tid)
{
    const int lgA = lgW + lgH, total = G.count << lgA;
    const int shift = 15 - 8 - (lgA >> 1);
    const int16_t *Cg = a.coeff + G.base;
    const uint16_t *Pg = a.pred + G.base;
    uint16_t *Og = a.recon + G.base;
    for (int e0 = 0; e0 < total; e0 += 2 * kRecThreads) {
        const int e = e0 + 2 * tid;
        const int b = min(e >> lgA, G.count - 1);
        const int o = __shfl(myRel, b) + (e & ((1 << lgA) - 1));
        if (e < total) {
            const v2s c = as_v2s(*reinterpret_cast<const uint32_t *>(Cg + o));
            const uint32_t p = *reinterpret_cast<const uint32_t *>(Pg + o);
            const int r0 = (int)(p & 0xFFFF) + (int16_t)(c[0] >> shift);
            const int r1 = (int)(p >> 16) + (int16_t)(c[1] >> shift);
            *reinterpret_cast<uint32_t *>(Og + o) = pack16(min(max(r0, 0), 1023), min(max(r1, 0), 1023));
        }
    }
}

// One workgroup = four waves; they walk groups of same-descriptor blocks.  A
// group issues its coefficient loads (rows below the zero-out only) and the
// basis fragments before its first wait; each stage's tiles are shared by the
// four waves.
// Six waves per SIMD: without the hint the compiler keeps MFMA accumulators in
// AGPRs and the unified register budget (75 + 8) allows five; asked for six it
// fits in 79 VGPRs with no spills (0.0784 -> 0.0720 ms on the 200k-TU
// workload; seven measured the same as six).
__global__ void __launch_bounds__(kRecThreads) __attribute__((amdgpu_waves_per_eu(6))) k_recon_mfma(RecArgs a, const RecGroup *__restrict__ groups,
                                                             const int32_t *__restrict__ rel)
{
    constexpr int NC = kRecStage / (8 * kRecThreads);     // 16-byte chunks per thread
    constexpr int PL = kRecPlane + kRecSlack;
    __shared__ uint4 stg[kRecThreads * NC];
    __shared__ uint64_t limb[2 * PL / 8];
    __shared__ int4 cstv[32];                               // offset terms: [0, 64) stage 1, [64, 128) stage 2
    int32_t *cst = reinterpret_cast<int32_t *>(cstv);
    const int16_t *S = reinterpret_cast<const int16_t *>(stg);
    uint8_t *PL0 = reinterpret_cast<uint8_t *>(limb), *PH = PL0 + PL;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    for (int gi = blockIdx.x; gi < a.n; gi += gridDim.x) {
        const RecGroup G = groups[gi];
        const int G_n = G.count;
        const int myRel = rel[G.first + min(lane, G_n - 1)];
        const uint32_t desc = G.desc;
        const int lgW = desc & 15, lgH = (desc >> 4) & 15;
        const int hor = (desc >> 8) & 15, ver = (desc >> 12) & 15;
        if (hor == kSKIP) {
            recon_skip(a, G, myRel, lgW, lgH, tid);
            continue;
        }
        const int W = 1 << lgW, H = 1 << lgH;
        const int nzW = tr_kept(hor, W), nzH = tr_kept(ver, H);
        const int lgNzW = 31 - __builtin_clz(nzW), lgNzH = 31 - __builtin_clz(nzH);
        const int XS = rec_stride(nzW);
        const int8_t *TH = a.mats + a.matOff[ver][lgH];
        const int8_t *TW = a.mats + a.matOff[hor][lgW];

        // ---- loads: basis fragments of both stages, then the coefficient rows ----
        // fragments past the block's side re-read row 0: the loads are
        // unconditional, so the waits below count them exactly
        long b1[4], a2[4];
#pragma unroll
        for (int mi = 0; mi < 4; mi++) {
            b1[mi] = load8(TH + (16 * mi < H ? 16 * mi + lr : lr) * 32 + 8 * lq);
            a2[mi] = load8(TW + (16 * mi < W ? 16 * mi + lr : lr) * 32 + 8 * lq);
        }
        // limb-offset terms of both bases (max(16, N) entries each) into LDS
        if (tid < 64) {
            const int32_t *CH = reinterpret_cast<const int32_t *>(a.mats + a.constOff[ver][lgH]);
            const int32_t *CW = reinterpret_cast<const int32_t *>(a.mats + a.constOff[hor][lgW]);
            cst[tid] = CH[min(tid, max(16, H) - 1)];
            cst[64 + tid] = CW[min(tid, max(16, W) - 1)];
        }
        const int16_t *Cg = a.coeff + G.base;
        if (W >= 8) {
            const int lgCpb = lgNzH + lgW - 3, nch = G_n << lgCpb;
            uint4 c[NC];
            stage_group<uint4, NC, kRecThreads>(c, Cg, myRel, lgCpb, nch, 8, tid);
            stage_store<uint4, NC, kRecThreads>(c, stg, tid);
        } else if (W == 4) {
            const int lgCpb = lgNzH, nch = G_n << lgCpb;
            uint2 c[2 * NC];
            stage_group<uint2, 2 * NC, kRecThreads>(c, Cg, myRel, lgCpb, nch, 4, tid);
            stage_store<uint2, 2 * NC, kRecThreads>(c, stg, tid);
        } else {
            const int lgCpb = lgNzH, nch = G_n << lgCpb;
            uint32_t c[4 * NC];
            stage_group<uint32_t, 4 * NC, kRecThreads>(c, Cg, myRel, lgCpb, nch, 2, tid);
            stage_store<uint32_t, 4 * NC, kRecThreads>(c, stg, tid);
        }
        __syncthreads();

        // ---- stage 1: M = columns (b, x < nzW), N = i < H, K = k < nzH ----
        const int cols1 = G_n << lgNzW;
        const Split s1 = work_split<kRecWaves>((cols1 + 15) >> 4, wave);
        for (int t = s1.t0; 16 * t < cols1; t += s1.tstep) {
            const int n0 = 16 * t;
            // A fragment: column m = n0 + lr, rows k = 8 lq .. +7 (clamped reads;
            // rows past nzH meet zero basis bytes)
            const int m = min(n0 + lr, cols1 - 1);
            const int16_t *col = S + ((m >> lgNzW) << (lgNzH + lgW)) + (m & (nzW - 1));
            uint32_t w[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k0 = min(8 * lq + 2 * q, nzH - 1), k1 = min(8 * lq + 2 * q + 1, nzH - 1);
                w[q] = pack16(col[k0 << lgW], col[k1 << lgW]);
            }
            long fl, fh;
            limbs2(w, fl, fh);
            // this lane's D: columns n0 + 4 lq + r (r = 0..3) of output row i
            const int oc = n0 + 4 * lq, ob = oc >> lgNzW, ox = oc & (nzW - 1);
#pragma unroll
            for (int mi = 0; mi < 4; mi++) {
                if (16 * mi >= H) break;
                if (!split_takes(s1, mi)) continue;
                // offset term of output row i and the rounding ride in dl
                const int c1 = cst[16 * mi + lr] + 64;                            // + stage-1 rounding
                v4i dl = {c1, c1, c1, c1}, dh = {0, 0, 0, 0};
                dl = __builtin_amdgcn_mfma_i32_16x16x32_i8(fl, b1[mi], dl, 0, 0, 0);
                dh = __builtin_amdgcn_mfma_i32_16x16x32_i8(fh, b1[mi], dh, 0, 0, 0);
                int e[4];
#pragma unroll
                for (int r = 0; r < 4; r++) e[r] = clip16((dl[r] + (dh[r] << 8)) >> 7);
                const uint32_t LL = bytes4(e[0], e[1], e[2], e[3]) ^ 0x80808080u;
                const uint32_t LH = bytes4(e[0] >> 8, e[1] >> 8, e[2] >> 8, e[3] >> 8);
                const int i = 16 * mi + lr;
                if (i < H && ob < G_n) {
                    const int o = ((ob << lgH) + i) * XS + ox;
                    if (nzW >= 4) {
                        *reinterpret_cast<uint32_t *>(PL0 + o) = LL;
                        *reinterpret_cast<uint32_t *>(PH + o) = LH;
                    } else {
                        // 2-wide blocks: columns r = 0, 1 belong to block ob, 2, 3 to block ob + 1
                        *reinterpret_cast<uint16_t *>(PL0 + o) = (uint16_t)LL;
                        *reinterpret_cast<uint16_t *>(PH + o) = (uint16_t)LH;
                        if (ob + 1 < G_n) {
                            const int o2 = o + (H * XS);
                            *reinterpret_cast<uint16_t *>(PL0 + o2) = (uint16_t)(LL >> 16);
                            *reinterpret_cast<uint16_t *>(PH + o2) = (uint16_t)(LH >> 16);
                        }
                    }
                }
            }
        }
        __syncthreads();

        // ---- stage 2: M = output column o < W, N = rows (b, i), K = x < nzW ----
        const int rows2 = G_n << lgH;
        const uint16_t *Pg = a.pred + G.base;
        uint16_t *Og = a.recon + G.base;
        const Split s2 = work_split<kRecWaves>((rows2 + 15) >> 4, wave);
        for (int t = s2.t0; 16 * t < rows2; t += s2.tstep) {
            const int n0 = 16 * t;
            const int row = min(n0 + lr, rows2 - 1);
            const int lrow = row * XS + 8 * lq;                // bytes past nzW meet zero basis bytes
            const long pl = *reinterpret_cast<const long *>(PL0 + lrow);
            const long ph = *reinterpret_cast<const long *>(PH + lrow);
            const int b = row >> lgH, i = row & (H - 1);
            const int rowOff = __shfl(myRel, b) + (i << lgW);
            const bool rowOk = n0 + lr < rows2;
#pragma unroll
            for (int mi = 0; mi < 4; mi++) {
                if (16 * mi >= W) break;
                if (!split_takes(s2, mi)) continue;
                // offset terms of output columns o0 .. o0 + 3 and the rounding ride in dl
                const int4 c2 = cstv[16 + 4 * mi + lq];
                v4i dl = {c2.x + 512, c2.y + 512, c2.z + 512, c2.w + 512}, dh = {0, 0, 0, 0};
                dl = __builtin_amdgcn_mfma_i32_16x16x32_i8(a2[mi], pl, dl, 0, 0, 0);
                dh = __builtin_amdgcn_mfma_i32_16x16x32_i8(a2[mi], ph, dh, 0, 0, 0);
                int v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) v[r] = clip16((dl[r] + (dh[r] << 8)) >> 10);
                const int o0 = 16 * mi + 4 * lq;                // this lane's output columns o0 .. o0 + 3
                if (rowOk && o0 < W) {
                    const int off = rowOff + o0;
                    if (W >= 4) {
                        const uint2 p = *reinterpret_cast<const uint2 *>(Pg + off);
                        const int q0 = min(max((int)(p.x & 0xFFFF) + v[0], 0), 1023);
                        const int q1 = min(max((int)(p.x >> 16) + v[1], 0), 1023);
                        const int q2 = min(max((int)(p.y & 0xFFFF) + v[2], 0), 1023);
                        const int q3 = min(max((int)(p.y >> 16) + v[3], 0), 1023);
                        *reinterpret_cast<uint2 *>(Og + off) = make_uint2(pack16(q0, q1), pack16(q2, q3));
                    } else {
                        const uint32_t p = *reinterpret_cast<const uint32_t *>(Pg + off);
                        const int q0 = min(max((int)(p & 0xFFFF) + v[0], 0), 1023);
                        const int q1 = min(max((int)(p >> 16) + v[1], 0), 1023);
                        *reinterpret_cast<uint32_t *>(Og + off) = pack16(q0, q1);
                    }
                }
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
struct RecPlan {
    RecGroup *dGroups = nullptr;
    int32_t *dRel = nullptr;
    int n = 0, cap = 0, relCap = 0;
    bool valid = false;
};
std::mutex g_rplan_mu;
std::vector<std::pair<xg_session *, RecPlan>> g_rplans;

RecPlan *rplan_for(xg_session *s, bool create)
{
    for (auto &p : g_rplans)
        if (p.first == s) return &p.second;
    if (!create) return nullptr;
    g_rplans.push_back({s, RecPlan()});
    return &g_rplans.back().second;
}

int32_t rplan_fail(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_rplan_mu);
    if (RecPlan *p = rplan_for(s, false)) p->valid = false;
    return XG_ERR_ARG;
}
}  // namespace

extern "C" {

int32_t xg_recon_plan(xg_session *s, const int64_t *offset, const uint32_t *desc, int32_t n)
{
    if (!s || n < 0 || (n > 0 && (!offset || !desc))) {
        set_error(XG_ERR_ARG, "xg_recon_plan: bad arguments");
        return s ? rplan_fail(s) : XG_ERR_ARG;
    }
    if (!rec_mats_init()) return XG_ERR_HIP;
    static const int8_t mts[6][2] = {{kDCT2, kDCT2}, {kSKIP, kSKIP}, {kDST7, kDST7},
                                     {kDCT8, kDST7}, {kDST7, kDCT8}, {kDCT8, kDCT8}};
    struct Item {
        long long off;
        uint32_t desc;
    };
    std::vector<Item> items((size_t)n);
    for (int i = 0; i < n; i++) {
        const uint32_t d = desc[i];
        const int lgW = d & 15, lgH = (d >> 4) & 15, m = (d >> 8) & 15, implicit = (d >> 12) & 1;
        if (lgW < 1 || lgW > 6 || lgH < 1 || lgH > 6 || m > 5 || (d >> 13) || offset[i] < 0) {
            set_error(XG_ERR_ARG, "xg_recon_plan: block %d has descriptor 0x%x", i, d);
            return rplan_fail(s);
        }
        int hor = mts[m][0], ver = mts[m][1];
        const int W = 1 << lgW, H = 1 << lgH;
        if (implicit && m == 0) {                      // h266_inverse_trans.c:149-159
            if (W >= 4 && W <= 16) hor = kDST7;
            if (H >= 4 && H <= 16) ver = kDST7;
        }
        if (hor != kSKIP && ((hor != kDCT2 && (W < 4 || W > 32)) || (ver != kDCT2 && (H < 4 || H > 32)))) {
            set_error(XG_ERR_ARG, "xg_recon_plan: block %d: DST-VII/DCT-VIII need sides 4..32", i);
            return rplan_fail(s);
        }
        if (offset[i] % std::min(8, W)) {
            set_error(XG_ERR_ARG, "xg_recon_plan: block %d: offset must be a multiple of min(8, width) samples", i);
            return rplan_fail(s);
        }
        items[i].off = offset[i];
        items[i].desc = (uint32_t)(lgW | (lgH << 4) | (hor << 8) | (ver << 12));
    }
    // largest blocks first, blocks of one descriptor in memory order
    std::sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
        const int ax = (x.desc & 15) + ((x.desc >> 4) & 15), ay = (y.desc & 15) + ((y.desc >> 4) & 15);
        if (ax != ay) return ax > ay;
        if (x.desc != y.desc) return x.desc < y.desc;
        return x.off < y.off;
    });
    std::vector<RecGroup> groups;
    std::vector<int32_t> rel((size_t)n);
    for (int i = 0; i < n;) {
        const uint32_t d = items[i].desc;
        const int cap = rec_group_cap(1 << (d & 15), 1 << ((d >> 4) & 15), (d >> 8) & 15, (d >> 12) & 15);
        RecGroup g;
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
    }
    const int ng = (int)groups.size();
    std::lock_guard<std::mutex> g(g_rplan_mu);
    RecPlan *p = rplan_for(s, true);
    p->valid = false;
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    if (!grow(p->dGroups, p->cap, std::max(ng, 1), st) || !grow(p->dRel, p->relCap, std::max(n, 1), st))
        return XG_ERR_NOMEM;
    if (ng && !XG_HIP(hipMemcpyAsync(p->dGroups, groups.data(), sizeof(RecGroup) * (size_t)ng, hipMemcpyHostToDevice,
                                     st)))
        return XG_ERR_HIP;
    if (n && !XG_HIP(hipMemcpyAsync(p->dRel, rel.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, st)))
        return XG_ERR_HIP;
    if (!XG_HIP(hipStreamSynchronize(st))) return XG_ERR_HIP;
    p->n = ng;
    p->valid = true;
    return XG_OK;
}

int32_t xg_recon_run(xg_session *s, const int16_t *coeff, const xg_pixel *pred, xg_pixel *recon)
{
    if (!s || !coeff || !pred || !recon) {
        set_error(XG_ERR_ARG, "xg_recon_run: bad arguments");
        return XG_ERR_ARG;
    }
    RecPlan *p;
    {
        std::lock_guard<std::mutex> g(g_rplan_mu);
        p = rplan_for(s, false);
    }
    if (!p || !p->valid) {
        set_error(XG_ERR_STATE, "xg_recon_run: no plan (call xg_recon_plan first; a failed plan leaves none)");
        return XG_ERR_STATE;
    }
    if (p->n == 0) return XG_OK;
    RecArgs a;
    a.coeff = coeff;
    a.pred = pred;
    a.recon = recon;
    a.n = p->n;
    a.mats = g_rmats;
    memcpy(a.matOff, g_rm.off, sizeof(a.matOff));
    memcpy(a.constOff, g_rm.coff, sizeof(a.constOff));
    static const int slots = resident_slots((const void *)k_recon_mfma, kRecThreads);
    hipStream_t st = (hipStream_t)xg_session_stream(s);
    int ev[2];
    const bool timed = session_timer_begin(s, 5, ev);
    k_recon_mfma<<<std::min(a.n, slots), kRecThreads, 0, st>>>(a, p->dGroups, p->dRel);
    if (timed) session_timer_end(s, 5, ev);
    return XG_HIP(hipGetLastError()) ? XG_OK : XG_ERR_HIP;
}

}  // extern "C"

void xg::recon_release(xg_session *s)
{
    std::lock_guard<std::mutex> g(g_rplan_mu);
    for (size_t i = 0; i < g_rplans.size(); i++)
        if (g_rplans[i].first == s) {
            if (g_rplans[i].second.dGroups) (void)hipFree(g_rplans[i].second.dGroups);
            if (g_rplans[i].second.dRel) (void)hipFree(g_rplans[i].second.dRel);
            g_rplans.erase(g_rplans.begin() + (long)i);
            return;
        }
}

// xg_quant.hpp -- the quantiser arithmetic shared by the batched quantiser
// (xg_quant.hip) and the forward transform's fused quantising epilogue
// (xg_fdct.hip).
//
// Per TU, the parameters Xin266Transform derives
// (h266_encoder/source/h266_trans_quant.c:110-120) from the quantParam table
// of Xin266ConstructQuantParam (h266_encoder/source/h266_encoder_create.c:
// 1677-1705; level scales :94-104), bit depth 10, from the descriptor
//   lgW | lgH << 4 | mtsIdx << 8 | lgCGW << 12 | lgCGH << 15 | qp << 18
//   | rdoq << 24 | intraFrame << 25
// and the per-coefficient arithmetic of Xin266QIQCodingGroup
// (h266_encoder/source/h266_quant_inv_quant.c:27-78), int32 as its SINT32
// expressions (wrapping).
#pragma once
#include <hip/hip_runtime.h>

namespace xg {

// qMult[adj][rem] and iqMult[adj][rem] of h266_encoder_create.c:94-104 as
// register selects (a per-lane table index would be a memory load): the
// odd-area set (adj = 1) is the base set shifted by half an octave --
// entry rem is base entry rem + 3, or base entry rem - 3 halved (forward) /
// doubled (inverse) -- which reproduces its twelve values exactly.
__device__ __forceinline__ void level_scales(int adj, int rem, int &q, int &iq)
{
    int i = rem, sh = 0;
    if (adj) {
        i = rem < 3 ? rem + 3 : rem - 3;
        sh = rem >= 3;
    }
    q = i == 0 ? 26214 : i == 1 ? 23302 : i == 2 ? 20560 : i == 3 ? 18396 : i == 4 ? 16384 : 14564;
    iq = i == 0 ? 40 : i == 1 ? 45 : i == 2 ? 51 : i == 3 ? 57 : i == 4 ? 64 : 72;
    q >>= sh;
    iq <<= sh;
}

struct QParams {
    int qMult, qAdd, qShift, iqMult, iqAdd, iqShift;
};

__device__ __forceinline__ QParams quant_params(uint32_t d)
{
    const int lgW = d & 15, lgH = (d >> 4) & 15, skip = ((d >> 8) & 15) == 1;
    const int qp = (d >> 18) & 63, rdoq = (d >> 24) & 1, intra = (d >> 25) & 1;
    const int lgTu = lgW + lgH, adj = (lgTu & 1) && !skip, lgSize = lgTu >> 1;
    const int per = qp / 6, rem = qp - 6 * per;
    QParams p;
    p.qShift = 14 + 5 + per - lgSize - adj;          // 14 + (15 - bit depth) + qp / 6
    p.iqShift = 6 - 5 + lgSize + adj;                // 6 - (15 - bit depth)
    int qs, iqs;
    level_scales(adj, rem, qs, iqs);
    p.qMult = qs;
    p.iqMult = iqs << per;
    p.qAdd = (rdoq && !skip) ? 1 << (p.qShift - 1) : (intra ? 171 : 85) << (p.qShift - 9);
    p.iqAdd = !skip ? 1 << (p.iqShift - 1) : 1 << ((p.iqShift - lgSize) + 5 - 1);
    return p;
}

// level of t; r receives its dequantised value
__device__ __forceinline__ int qiq(int t, const QParams &p, int &r)
{
    const int a = t < 0 ? -t : t;
    int q = (int)((unsigned)a * (unsigned)p.qMult + (unsigned)p.qAdd) >> p.qShift;
    q = min(max(t < 0 ? -q : q, -32768), 32767);
    r = min(max((int)((unsigned)q * (unsigned)p.iqMult + (unsigned)p.iqAdd) >> p.iqShift, -32768), 32767);
    return q;
}

// qiq for |t| < 2^24 (forward-transform outputs: below 2^22 for any int16
// residual), with the full-rate 24-bit multiply: its low 32 bits are those of
// the wrapping int32 product, so the result is qiq's.
__device__ __forceinline__ int qiq24(int t, const QParams &p, int &r)
{
    const unsigned a = (unsigned)(t < 0 ? -t : t);
    int q = (int)(__umul24(a, (unsigned)p.qMult) + (unsigned)p.qAdd) >> p.qShift;
    q = min(max(t < 0 ? -q : q, -32768), 32767);
    r = min(max((int)((unsigned)q * (unsigned)p.iqMult + (unsigned)p.iqAdd) >> p.iqShift, -32768), 32767);
    return q;
}

}  // namespace xg

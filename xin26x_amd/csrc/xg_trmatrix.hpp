// xg_trmatrix.hpp -- VVC forward-transform basis matrices, built at compile
// time.  ITU-T H.266 clause 8.7.4: DCT-II of size N uses rows k*64/N of the
// 64-point matrix; DST-VII N in {4,8,16,32} from its sine-magnitude table;
// DCT-VIII[k][n] = (-1)^k * DST-VII[k][N-1-n].  All entries fit int8.
#pragma once
#include <stdint.h>

namespace xg {

struct TrTables {
    int8_t dct64[64][64];
    int8_t dst4[4][4], dst8[8][8], dst16[16][16], dst32[32][32];
};

namespace trdetail {
// |cos((2i+1) pi / (2L))| magnitudes of each dyadic level L (i < L/2), and
// |sin(j pi / (2N+1))| magnitudes of DST-VII (j = 1..N), as tabulated in H.266.
constexpr int cos64[32] = {91, 90, 90, 90, 88, 87, 86, 84, 83, 81, 79, 77, 73, 71, 69, 65,
                           62, 59, 56, 52, 48, 44, 41, 37, 33, 28, 24, 20, 15, 11, 7, 2};
constexpr int cos32[16] = {90, 90, 88, 85, 82, 78, 73, 67, 61, 54, 46, 38, 31, 22, 13, 4};
constexpr int cos16[8] = {90, 87, 80, 70, 57, 43, 25, 9};
constexpr int cos8[4] = {89, 75, 50, 18};
constexpr int cos4[2] = {83, 36};
constexpr int sin4[4] = {29, 55, 74, 84};
constexpr int sin8[8] = {17, 32, 46, 60, 71, 78, 85, 86};
constexpr int sin16[16] = {8, 17, 25, 33, 40, 48, 55, 62, 68, 73, 77, 81, 85, 87, 88, 88};
constexpr int sin32[32] = {4, 9, 13, 17, 21, 26, 30, 34, 38, 42, 46, 50, 53, 56, 60, 63,
                           66, 68, 72, 74, 77, 78, 80, 82, 84, 85, 86, 87, 88, 89, 90, 90};

constexpr int level_mag(int L, int i)
{
    return L == 64 ? cos64[i] : L == 32 ? cos32[i] : L == 16 ? cos16[i] : L == 8 ? cos8[i]
         : L == 4 ? cos4[i] : 64;
}

// Sign and magnitude of cos((2n+1) k pi / 128) scaled as in H.266.
constexpr int dct64_entry(int k, int n)
{
    if (k == 0) return 64;
    int odd = k, L = 64;
    while ((odd & 1) == 0) { odd >>= 1; L >>= 1; }
    const int m = ((2 * n + 1) * odd) % (4 * L);
    const int quadrant = m / L;
    if (quadrant == 0) return level_mag(L, (m - 1) / 2);
    if (quadrant == 1) return -level_mag(L, (2 * L - m - 1) / 2);
    if (quadrant == 2) return -level_mag(L, (m - 2 * L - 1) / 2);
    return level_mag(L, (4 * L - m - 1) / 2);
}

constexpr int sin_mag(int N, int j)
{
    return N == 4 ? sin4[j] : N == 8 ? sin8[j] : N == 16 ? sin16[j] : sin32[j];
}

// Sign and magnitude of sin((2k+1)(n+1) pi / (2N+1)).
constexpr int dst7_entry(int N, int k, int n)
{
    const int P = 2 * N + 1;
    int m = ((2 * k + 1) * (n + 1)) % (2 * P);
    if (m == 0 || m == P) return 0;
    int sign = 1;
    if (m > P) { m -= P; sign = -1; }
    const int j = m <= N ? m : P - m;
    return sign * sin_mag(N, j - 1);
}
}  // namespace trdetail

constexpr TrTables make_tr_tables()
{
    TrTables t{};
    for (int k = 0; k < 64; k++)
        for (int n = 0; n < 64; n++) t.dct64[k][n] = (int8_t)trdetail::dct64_entry(k, n);
    for (int k = 0; k < 4; k++)
        for (int n = 0; n < 4; n++) t.dst4[k][n] = (int8_t)trdetail::dst7_entry(4, k, n);
    for (int k = 0; k < 8; k++)
        for (int n = 0; n < 8; n++) t.dst8[k][n] = (int8_t)trdetail::dst7_entry(8, k, n);
    for (int k = 0; k < 16; k++)
        for (int n = 0; n < 16; n++) t.dst16[k][n] = (int8_t)trdetail::dst7_entry(16, k, n);
    for (int k = 0; k < 32; k++)
        for (int n = 0; n < 32; n++) t.dst32[k][n] = (int8_t)trdetail::dst7_entry(32, k, n);
    return t;
}

constexpr TrTables kTr = make_tr_tables();

enum TrType { kDCT2 = 0, kDCT8 = 1, kDST7 = 2, kSKIP = 4 };

// Entry (k, n) of the N-point matrix of `type` (host and device).
__host__ __device__ inline int tr_entry(const TrTables &t, int type, int N, int k, int n)
{
    if (type == kDCT2) return t.dct64[k * (64 / N)][n];
    const int nn = type == kDST7 ? n : N - 1 - n;
    const int sign = (type == kDCT8 && (k & 1)) ? -1 : 1;
    int v = N == 4 ? t.dst4[k][nn] : N == 8 ? t.dst8[k][nn] : N == 16 ? t.dst16[k][nn] : t.dst32[k][nn];
    return sign * v;
}

// Outputs a 1-D forward pass keeps before the VVC zero-out.
__host__ __device__ inline int tr_kept(int type, int N)
{
    if (type == kDCT2) return N == 64 ? 32 : N;
    return N == 32 ? 16 : N;
}

}  // namespace xg

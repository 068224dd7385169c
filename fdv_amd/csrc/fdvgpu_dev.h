// fdvgpu_dev.h -- device helpers shared by the analysis and motion-table kernels:
// scan / quantiser tables, wave reductions, and quarter-pel interpolation from
// LDS windows (reference: yhbsh/fdv include/fdv.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Workgroups go to the 8 XCDs round-robin by linear id, so with a grid width
// that is a multiple of 8, blockIdx.x % 8 names the XCD.  A one-workgroup-per-MB
// kernel gives each XCD one vertical strip of ceil(mw / 8) MB columns (MBs k =
// blockIdx.x / 8 of the strip in raster order): the reference windows that
// neighbouring MBs share are then fetched into one L2.  A grid width of
// xcd_grid(mw, mh) covers every MB; false for the workgroups beyond it.
__host__ __device__ __forceinline__ int xcd_grid(int mw, int mh) { return mh * ((mw + 7) >> 3) * 8; }
__device__ __forceinline__ bool xcd_mb(int bx, int mw, int mh, int &i) {
    const int sw = (mw + 7) >> 3, k = bx >> 3;
    const int row = k / sw, col = (bx & 7) * sw + k % sw;
    if (row >= mh || col >= mw) return false;
    i = row * mw + col;
    return true;
}

static __device__ __constant__ uint8_t k_SCAN4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
static __device__ __constant__ uint8_t k_INV4[16] = {0, 1, 5, 6, 2, 4, 7, 12, 3, 8, 11, 13, 9, 10, 14, 15};
static __device__ __constant__ uint8_t k_INV8[64] = {
     0,  1,  5,  6, 14, 15, 27, 28,  2,  4,  7, 13, 16, 26, 29, 42,
     3,  8, 12, 17, 25, 30, 41, 43,  9, 11, 18, 24, 31, 40, 44, 53,
    10, 19, 23, 32, 39, 45, 52, 54, 20, 22, 33, 38, 46, 51, 55, 60,
    21, 34, 37, 47, 50, 56, 59, 61, 35, 36, 48, 49, 57, 58, 62, 63};
static __device__ __constant__ int k_MF[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490}, {10082, 4194, 6554},
                                          {9362, 3647, 5825},  {8192, 3355, 5243},  {7282, 2893, 4559}};
static __device__ __constant__ int k_DV[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16},
                                          {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
static __device__ __constant__ int8_t k_LDSP[8][2] = {{0, 8}, {0, -8}, {8, 0}, {-8, 0}, {4, 4}, {4, -4}, {-4, 4}, {-4, -4}};
static __device__ __constant__ int8_t k_SDSP[4][2] = {{0, 4}, {0, -4}, {4, 0}, {-4, 0}};
static __device__ __constant__ int8_t k_CROSS[4][2] = {{0, -1}, {0, 1}, {-1, 0}, {1, 0}};

#define WIN 64        // luma window side, origin (bx-24, by-24)
#define WOFF 24
#define CWIN 32       // chroma window side, origin (cbx-12, cby-12)
#define CWOFF 12

// Exact floors of the coded modes, in bits (fdv.h:5887-5896; mode 3, flag 2).
#define FL_INTRA 14
#define FL_INTER16 16
#define FL_INTER8 19

// ------------------------------------------------------------------ helpers

// A pointer read out of a struct in memory is generic to the compiler, so its
// loads become flat instructions, which also count against the LDS counter:
// every later LDS wait then waits for them too. Device buffers handed in
// that way are global memory; this says so.
template <class T> using gmem_t = __attribute__((address_space(1))) T;
template <class T> __device__ __forceinline__ const gmem_t<T> *gmem(const T *p) {
    return (const gmem_t<T> *)p;
}
template <class T> __device__ __forceinline__ gmem_t<T> *gmemw(T *p) { return (gmem_t<T> *)p; }
__device__ __forceinline__ int clip8(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
__device__ __forceinline__ uint32_t zmap(int v) { return ((uint32_t)v << 1) ^ (uint32_t)(v >> 31); }
__device__ __forceinline__ int egbits(uint32_t v) { return 2 * (31 - __clz((int)(v + 1u))) + 1; }
__device__ __forceinline__ int leblen(uint32_t u) {
    return u < (1u << 7) ? 1 : u < (1u << 14) ? 2 : u < (1u << 21) ? 3 : u < (1u << 28) ? 4 : 5;
}
__device__ __forceinline__ int pclass(int i) { int r = (i >> 2) & 1, c = i & 1; return r == c ? r : 2; }

template <int W> __device__ __forceinline__ int gsum(int v) {
#pragma unroll
    for (int o = 1; o < W; o <<= 1) v += __shfl_xor(v, o);
    return v;
}
template <int W> __device__ __forceinline__ int gmax(int v) {
#pragma unroll
    for (int o = 1; o < W; o <<= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
template <int W> __device__ __forceinline__ int gor(int v) {
#pragma unroll
    for (int o = 1; o < W; o <<= 1) v |= __shfl_xor(v, o);
    return v;
}

// ---- luma / chroma interpolation from LDS windows (H.264 fig. 8-4 naming;
//      fdv.h:2931-2990 and the block form 3082-3230 it equals)
__device__ __forceinline__ int tap6(int a, int b, int c, int d, int e, int f) {
    return a - 5 * b + 20 * c + 20 * d - 5 * e + f;
}
// Planes of row stride S (an LDS window of WIN, or a tile).
template <int S = WIN> __device__ __forceinline__ int wpx(const uint8_t *w, int x, int y) { return w[y * S + x]; }
template <int S = WIN> __device__ __forceinline__ int whr(const uint8_t *w, int x, int y) {
    const uint8_t *p = w + y * S + x;
    return tap6(p[-2], p[-1], p[0], p[1], p[2], p[3]);
}
template <int S = WIN> __device__ __forceinline__ int wvr(const uint8_t *w, int x, int y) {
    const uint8_t *p = w + y * S + x;
    return tap6(p[-2 * S], p[-S], p[0], p[S], p[2 * S], p[3 * S]);
}
template <int S = WIN> __device__ __forceinline__ int wb(const uint8_t *w, int x, int y) {
    return clip8((whr<S>(w, x, y) + 16) >> 5);
}
template <int S = WIN> __device__ __forceinline__ int wh(const uint8_t *w, int x, int y) {
    return clip8((wvr<S>(w, x, y) + 16) >> 5);
}
template <int S = WIN> __device__ __forceinline__ int wj(const uint8_t *w, int x, int y) {
    int v = tap6(whr<S>(w, x, y - 2), whr<S>(w, x, y - 1), whr<S>(w, x, y), whr<S>(w, x, y + 1),
                 whr<S>(w, x, y + 2), whr<S>(w, x, y + 3));
    return clip8((v + 512) >> 10);
}
__device__ __forceinline__ int av2(int a, int b) { return (a + b + 1) >> 1; }
__device__ __forceinline__ int pick8(int i, int v0, int v1, int v2, int v3, int v4, int v5, int v6, int v7) {
    int r = v0;
    r = i == 1 ? v1 : r;
    r = i == 2 ? v2 : r;
    r = i == 3 ? v3 : r;
    r = i == 4 ? v4 : r;
    r = i == 5 ? v5 : r;
    r = i == 6 ? v6 : r;
    return i == 7 ? v7 : r;
}
// Error of 4 vertically adjacent predicted luma samples: the samples at
// (X, Y .. Y + 3) of plane w (row stride S, 4-byte aligned, S a multiple of 4)
// at quarter-pel phase ph, against cur[0], cur[cs], cur[2 cs], cur[3 cs]; the
// sum of |d| (SQ false) or d^2 (SQ true).  The 9 x 6 samples those need
// (rows Y - 2 .. Y + 6, columns X - 2 .. X + 3) are read once as 27 dwords, and
// every quarter-pel sample of lumaq's definition is formed in registers; the
// phase may differ between lanes, so the sample is picked without branches as
// av2 of two of the eight candidate values.  Reads dwords up to 9 bytes past
// column X + 3 of each row.  BURST issues all 27 reads before any is used (for
// a caller on a serial path with registers to spare; otherwise the scheduler
// may interleave them with their uses, one read in flight at a time).
template <int S, bool SQ, bool BURST = false>
__device__ __forceinline__ int qerr4(const uint8_t *w, int X, int Y, int ph, const uint8_t *cur, int cs) {
    const int c0 = X - 2, sh = c0 & 3;
    uint32_t R[9][3];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const uint32_t *row = (const uint32_t *)(w + (Y - 2 + i) * S + (c0 & ~3));
        R[i][0] = row[0]; R[i][1] = row[1]; R[i][2] = row[2];
    }
    if (BURST) __builtin_amdgcn_sched_barrier(0);
    int V[9][6];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const uint32_t w0 = R[i][0], w1 = R[i][1], w2 = R[i][2];
        const uint32_t a = __builtin_amdgcn_alignbyte(w1, w0, sh), b = __builtin_amdgcn_alignbyte(w2, w1, sh);
#pragma unroll
        for (int j = 0; j < 4; ++j) V[i][j] = (int)((a >> (8 * j)) & 255u);
        V[i][4] = (int)(b & 255u);
        V[i][5] = (int)((b >> 8) & 255u);
    }
    int Hr[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Hr[i] = tap6(V[i][0], V[i][1], V[i][2], V[i][3], V[i][4], V[i][5]);
    // per phase (4 bits each, phase 0 lowest): the two values averaged, as
    // indices into G, b, hh, j, G(x+1), G(y+1), b(y+1), hh(x+1)
    const uint64_t A = 0x7322332211101100ull, B = 0x6665733273224110ull;
    const int sa = (int)((A >> (4 * ph)) & 15u), sb = (int)((B >> (4 * ph)) & 15u);
    int acc = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int G = V[k + 2][2], Gr = V[k + 2][3], Gd = V[k + 3][2];
        const int b = clip8((Hr[k + 2] + 16) >> 5), bd = clip8((Hr[k + 3] + 16) >> 5);
        const int hh = clip8((tap6(V[k][2], V[k + 1][2], V[k + 2][2], V[k + 3][2], V[k + 4][2], V[k + 5][2]) + 16) >> 5);
        const int hhr = clip8((tap6(V[k][3], V[k + 1][3], V[k + 2][3], V[k + 3][3], V[k + 4][3], V[k + 5][3]) + 16) >> 5);
        const int jj = clip8((tap6(Hr[k], Hr[k + 1], Hr[k + 2], Hr[k + 3], Hr[k + 4], Hr[k + 5]) + 512) >> 10);
        const int d = cur[k * cs] - av2(pick8(sa, G, b, hh, jj, Gr, Gd, bd, hhr), pick8(sb, G, b, hh, jj, Gr, Gd, bd, hhr));
        acc += SQ ? d * d : (d < 0 ? -d : d);
    }
    return acc;
}
// Predicted luma sample whose whole-pel position is (X, Y) of a plane of row
// stride S, at quarter-pel phase ph = (mvy & 3) << 2 | (mvx & 3).
template <int S>
__device__ __forceinline__ int lumaq_at(const uint8_t *w, int X, int Y, int ph) {
    switch (ph) {
    case 0:  return wpx<S>(w, X, Y);
    case 1:  return av2(wpx<S>(w, X, Y), wb<S>(w, X, Y));
    case 2:  return wb<S>(w, X, Y);
    case 3:  return av2(wb<S>(w, X, Y), wpx<S>(w, X + 1, Y));
    case 4:  return av2(wpx<S>(w, X, Y), wh<S>(w, X, Y));
    case 5:  return av2(wb<S>(w, X, Y), wh<S>(w, X, Y));
    case 6:  return av2(wb<S>(w, X, Y), wj<S>(w, X, Y));
    case 7:  return av2(wb<S>(w, X, Y), wh<S>(w, X + 1, Y));
    case 8:  return wh<S>(w, X, Y);
    case 9:  return av2(wh<S>(w, X, Y), wj<S>(w, X, Y));
    case 10: return wj<S>(w, X, Y);
    case 11: return av2(wj<S>(w, X, Y), wh<S>(w, X + 1, Y));
    case 12: return av2(wh<S>(w, X, Y), wpx<S>(w, X, Y + 1));
    case 13: return av2(wh<S>(w, X, Y), wb<S>(w, X, Y + 1));
    case 14: return av2(wj<S>(w, X, Y), wb<S>(w, X, Y + 1));
    default: return av2(wh<S>(w, X + 1, Y), wb<S>(w, X, Y + 1));
    }
}
// Predicted luma sample for the MB-relative pixel (x, y) displaced by (mvx, mvy).
__device__ __forceinline__ int lumaq(const uint8_t *w, int x, int y, int mvx, int mvy) {
    return lumaq_at<WIN>(w, WOFF + x + (mvx >> 2), WOFF + y + (mvy >> 2), ((mvy & 3) << 2) | (mvx & 3));
}
// Bilinear chroma sample for chroma-MB-relative (x, y) with a chroma quarter-pel
// vector (fdv.h:3260-3283).
__device__ __forceinline__ int chromaq(const uint8_t *w, int x, int y, int mvx, int mvy) {
    int X = CWOFF + x + (mvx >> 2), Y = CWOFF + y + (mvy >> 2), fx = mvx & 3, fy = mvy & 3;
    const uint8_t *p = w + Y * CWIN + X;
    int top = p[0] * (4 - fx) + p[1] * fx;
    int bot = p[CWIN] * (4 - fx) + p[CWIN + 1] * fx;
    return (top * (4 - fy) + bot * fy + 8) >> 4;
}


// ---- cross-lane reductions with DPP (no LDS round trip).  On gfx950 row_ror:N
//      gives lane i the value of lane (i - N) mod 16 of its 16-lane row.
template <int CTRL> __device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
#define DPP_XOR1 0xB1     // quad_perm [1,0,3,2]
#define DPP_XOR2 0x4E     // quad_perm [2,3,0,1]
#define DPP_ROR4 0x124
#define DPP_ROR8 0x128
// every lane of each 16-lane row receives the row's sum / max
__device__ __forceinline__ int row_sum16(int v) {
    v += dpp_mov<DPP_XOR1>(v); v += dpp_mov<DPP_XOR2>(v);
    v += dpp_mov<DPP_ROR4>(v); v += dpp_mov<DPP_ROR8>(v);
    return v;
}
__device__ __forceinline__ int row_max16(int v) {
    v = max(v, dpp_mov<DPP_XOR1>(v)); v = max(v, dpp_mov<DPP_XOR2>(v));
    v = max(v, dpp_mov<DPP_ROR4>(v)); v = max(v, dpp_mov<DPP_ROR8>(v));
    return v;
}
__device__ __forceinline__ int row_min16(int v) {
    v = min(v, dpp_mov<DPP_XOR1>(v)); v = min(v, dpp_mov<DPP_XOR2>(v));
    v = min(v, dpp_mov<DPP_ROR4>(v)); v = min(v, dpp_mov<DPP_ROR8>(v));
    return v;
}
// inclusive prefix sum over each 16-lane row (row_shr: lanes with no source
// lane in the row read 0)
#define DPP_SHR1 0x111
#define DPP_SHR2 0x112
#define DPP_SHR4 0x114
#define DPP_SHR8 0x118
__device__ __forceinline__ int row_prefix16(int v) {
    v += dpp_mov<DPP_SHR1>(v); v += dpp_mov<DPP_SHR2>(v);
    v += dpp_mov<DPP_SHR4>(v); v += dpp_mov<DPP_SHR8>(v);
    return v;
}
// sum over each aligned group of 8 lanes; valid in lanes 8g+4 .. 8g+7
__device__ __forceinline__ int oct_sum8(int v) {
    v += dpp_mov<DPP_XOR1>(v); v += dpp_mov<DPP_XOR2>(v);
    v += dpp_mov<DPP_ROR4>(v);
    return v;
}
// wave-uniform totals
__device__ __forceinline__ int wave_sum(int v) {
    v = row_sum16(v);
    return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) +
           __builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48);
}
__device__ __forceinline__ int wave_max(int v) {
    v = row_max16(v);
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
// order this wave's LDS writes before its later LDS reads by other lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Scan tables as constant expressions, for indexing register arrays in fully
// unrolled loops (fdv.h:1699).
__device__ __forceinline__ constexpr int k_SCAN4c(int k) {
    return k == 0 ? 0 : k == 1 ? 1 : k == 2 ? 4 : k == 3 ? 8 : k == 4 ? 5 : k == 5 ? 2 : k == 6 ? 3 :
           k == 7 ? 6 : k == 8 ? 9 : k == 9 ? 12 : k == 10 ? 13 : k == 11 ? 10 : k == 12 ? 7 :
           k == 13 ? 11 : k == 14 ? 14 : 15;
}
__device__ __forceinline__ constexpr int k_INV4c(int i) {
    return i == 0 ? 0 : i == 1 ? 1 : i == 2 ? 5 : i == 3 ? 6 : i == 4 ? 2 : i == 5 ? 4 : i == 6 ? 7 :
           i == 7 ? 12 : i == 8 ? 3 : i == 9 ? 8 : i == 10 ? 11 : i == 11 ? 13 : i == 12 ? 9 :
           i == 13 ? 10 : i == 14 ? 14 : 15;
}

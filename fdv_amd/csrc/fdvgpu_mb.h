// fdvgpu_mb.h -- per-MB device helpers shared by the encoder's kernels and the
// decoder's: intra 16x16 / 8x8 prediction with its SATD screen, and the intra
// handoff words through which an intra MB's intra neighbours read its edges
// (reference: yhbsh/fdv include/fdv.h).
#pragma once
#include "fdvgpu_dev.h"

// ---- intra 16x16 / 8x8 prediction (fdv.h:2851-2911).
struct Nb { int ht, hl, dc, a, b, c; };
__device__ __forceinline__ Nb nb_setup(const uint8_t *top, const uint8_t *left, int tl, int n, int ht, int hl) {
    Nb q; q.ht = ht; q.hl = hl;
    int s = 0;
    if (ht && hl) { for (int k = 0; k < n; ++k) s += top[k] + left[k]; q.dc = (s + n) / (2 * n); }
    else if (ht)  { for (int k = 0; k < n; ++k) s += top[k]; q.dc = (s + n / 2) / n; }
    else if (hl)  { for (int k = 0; k < n; ++k) s += left[k]; q.dc = (s + n / 2) / n; }
    else q.dc = 128;
    q.a = q.b = q.c = 0;
    if (ht && hl) {
        int hn = n / 2, H = 0, V = 0;
        for (int k = 1; k <= hn; ++k) {
            int j = hn - 1 - k;
            H += k * (top[hn - 1 + k] - (j < 0 ? tl : top[j]));
            V += k * (left[hn - 1 + k] - (j < 0 ? tl : left[j]));
        }
        q.a = 16 * (left[n - 1] + top[n - 1]);
        if (n == 16) { q.b = (5 * H + 32) >> 6; q.c = (5 * V + 32) >> 6; }
        else         { q.b = (17 * H + 16) >> 5; q.c = (17 * V + 16) >> 5; }
    }
    return q;
}
__device__ __forceinline__ bool nn_ok(int m, const Nb &q) {
    return m == 0 || (m == 1 && q.ht) || (m == 2 && q.hl) || (m == 3 && q.ht && q.hl);
}
__device__ __forceinline__ int ipred(int m, const Nb &q, const uint8_t *top, const uint8_t *left,
                                     int n, int x, int y) {
    if (m == 3 && q.ht && q.hl) {
        int c0 = n / 2 - 1;
        return clip8((q.a + q.c * (y - c0) + 16 + q.b * (x - c0)) >> 5);
    }
    if (m == 1 && q.ht) return top[x];
    if (m == 2 && q.hl) return left[y];
    return q.dc;
}

// SATD of the four 16x16 intra sub-modes (mb_satd, fdv.h:4411): one wavefront,
// lane = (mode, 4x4 block).  Returns mode m's total in sat[m] (uniform).
__device__ __forceinline__ void satd_modes(const uint8_t *cur, const Nb &nb, const uint8_t *top,
                                           const uint8_t *left, int *sat) {
    const int L = threadIdx.x & 63, m = L >> 4, b = L & 15, oy = (b >> 2) * 4, ox = (b & 3) * 4;
    int d[16], t[16], sa = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        d[k] = cur[(oy + (k >> 2)) * 16 + ox + (k & 3)] - ipred(m, nb, top, left, 16, ox + (k & 3), oy + (k >> 2));
#pragma unroll
    for (int k = 0; k < 4; ++k) {                     // rows (fdv.h:3803-3822)
        const int p0 = d[4 * k] + d[4 * k + 1], m0 = d[4 * k] - d[4 * k + 1];
        const int p1 = d[4 * k + 2] + d[4 * k + 3], m1 = d[4 * k + 2] - d[4 * k + 3];
        t[4 * k] = p0 + p1; t[4 * k + 1] = m0 + m1; t[4 * k + 2] = p0 - p1; t[4 * k + 3] = m0 - m1;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {                     // columns
        const int p0 = t[k] + t[4 + k], m0 = t[k] - t[4 + k];
        const int p1 = t[8 + k] + t[12 + k], m1 = t[8 + k] - t[12 + k];
        sa += abs(p0 + p1) + abs(m0 + m1) + abs(p0 - p1) + abs(m0 - m1);
    }
    sa = row_sum16(sa);
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) sat[mm] = __builtin_amdgcn_readlane(sa, 16 * mm);
}


// Intra handoff: what an intra MB's intra neighbours read of it, its bottom
// row and right column (luma 16 + 16, U and V 8 + 8 each), as 16 64-bit words
// of 4 payload bytes and the 32-bit frame-step id.  Relaxed agent-scope atomics
// keep them coherent across XCDs without cache-wide fences; the reader spins on
// the words it needs until their id is this step's (bounded, error flag).
//   words 0-3 luma bottom row, 4-5 U bottom row, 6-7 V bottom row,
//   8-11 luma right column, 12-13 U right column, 14-15 V right column.
#define IH_WORDS 16
__device__ __forceinline__ void ih_publish(uint64_t *ih, const uint8_t *yrec, const uint8_t *crec, unsigned step) {
    const int w = threadIdx.x;
    if (w >= IH_WORDS) return;
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int b;
        if (w < 4) b = yrec[15 * 16 + 4 * w + j];
        else if (w < 8) b = crec[(w >= 6 ? 64 : 0) + 7 * 8 + 4 * ((w - 4) & 1) + j];
        else if (w < 12) b = yrec[(4 * (w - 8) + j) * 16 + 15];
        else b = crec[(w >= 14 ? 64 : 0) + (4 * ((w - 12) & 1) + j) * 8 + 7];
        v |= (uint32_t)b << (8 * j);
    }
    __hip_atomic_store(ih + w, ((uint64_t)step << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ih_byte(const uint64_t *ih, int w, int byte, unsigned step, unsigned *ctl) {
    uint64_t v = __hip_atomic_load(ih + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (unsigned spins = 0; (unsigned)(v >> 32) != step; ++spins) {
        if (spins > (1u << 24)) { atomicOr(ctl + 1, 1u); break; }
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(ih + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return (int)((v >> (8 * byte)) & 0xff);
}


// fdvgpu_resid.h -- wave-local residual coding of one 8x8 region (one wavefront).
//
// The reference codes an n x n residual region by region (fdv_code_residual_inner,
// fdv.h:5300-5414): option A is four 4x4 transforms with dead-zone quantisation
// and RDOQ (code_block4 5214, fdv_rdoq4x4_inner 2493), option B one 8x8 transform
// without RDOQ (2614-2685); the cheaper in D + lambda*R wins, ties to A.
//
// Mapping on a 64-lane wavefront: the four 4x4 blocks take 16 lanes each.  Every
// lane holds its block's residual, coefficients and levels in registers (the
// 4x4 transforms are unrolled, no cross-lane traffic), and scores one RDOQ
// truncation point: lane c takes L = c - 1 (lane 0 also L = 15), so all
// candidates of all four blocks are scored at once.  One lane per block then
// applies the reference's sequential selection rule to the scores.  The
// latency form (LAT, for a caller on a serial path) scores L = c in lane c and
// L = -1 in closed form, takes the block's bit counts from DPP sums over the
// lanes' own coefficients, and selects by DPP reductions.  Option B runs with
// one lane per 8x8 coefficient, transposing through LDS.  Sums and maxima use
// DPP and readlane.
#pragma once
#include "fdvgpu_dev.h"

// Each source file that includes this has its own copy of the tables (constant
// memory belongs to a code object); each fills its copy at context creation.
static __constant__ int32_t c_T8[64];      // 8x8 DCT basis (built on the host from libm)
static __constant__ int32_t c_QS8[52];     // 8x8 quantiser steps

struct RegionWs {
    double rdJ[4][4][17];        // [region][block][L + 1] RDOQ scores
    int16_t levA[4][64];         // option A levels, four raster 4x4 blocks
    uint8_t recA[4][64];         // option A reconstruction, raster 8x8
    int32_t x32[4][64];
    int64_t x64[4][64];
    int32_t e[4], b[4];          // per region: SSD and bits of the chosen option
};

// ---- 4x4 transforms in registers (fdv.h:2233-2292)
__device__ __forceinline__ void fdct4_reg(const int *x, int *y) {
    int t[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int a = x[4 * i], b = x[4 * i + 1], c = x[4 * i + 2], d = x[4 * i + 3];
        const int s03 = a + d, d03 = a - d, s12 = b + c, d12 = b - c;
        t[4 * i] = s03 + s12; t[4 * i + 1] = 2 * d03 + d12; t[4 * i + 2] = s03 - s12; t[4 * i + 3] = d03 - 2 * d12;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int a = t[j], b = t[4 + j], c = t[8 + j], d = t[12 + j];
        const int s03 = a + d, d03 = a - d, s12 = b + c, d12 = b - c;
        y[j] = s03 + s12; y[4 + j] = 2 * d03 + d12; y[8 + j] = s03 - s12; y[12 + j] = d03 - 2 * d12;
    }
}
__device__ __forceinline__ void idct4_reg(const int *x, int *y) {
    int t[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int a = x[4 * i], b = x[4 * i + 1], c = x[4 * i + 2], d = x[4 * i + 3];
        const int e0 = a + c, e1 = a - c, e2 = (b >> 1) - d, e3 = b + (d >> 1);
        t[4 * i] = e0 + e3; t[4 * i + 1] = e1 + e2; t[4 * i + 2] = e1 - e2; t[4 * i + 3] = e0 - e3;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int a = t[j], b = t[4 + j], c = t[8 + j], d = t[12 + j];
        const int e0 = a + c, e1 = a - c, e2 = (b >> 1) - d, e3 = b + (d >> 1);
        y[j] = (e0 + e3 + 32) >> 6; y[4 + j] = (e1 + e2 + 32) >> 6;
        y[8 + j] = (e1 - e2 + 32) >> 6; y[12 + j] = (e0 - e3 + 32) >> 6;
    }
}
// position class of raster index i: 0 even/even, 1 odd/odd, 2 mixed (fdv.h:2423)
__device__ __forceinline__ constexpr int cls4(int i) {
    return (((i >> 2) & 1) == (i & 1)) ? ((i >> 2) & 1) : 2;
}
__device__ __forceinline__ uint32_t lds_word(const uint8_t *p) { return *(const uint32_t *)p; }

// The 8x8 basis and inverse zigzag as LDS copies, for a caller on a serial
// path: the option-B transform reads the basis at lane-dependent positions,
// which from constant memory are vector memory loads.
struct Tab8 {
    int32_t t8[64];
    uint8_t inv8[64];
};
__device__ __forceinline__ void tab8_fill(Tab8 &T, int t) {
    if (t < 64) { T.t8[t] = c_T8[t]; T.inv8[t] = k_INV8[t]; }
}

// One step of the RDOQ choice over a 16-lane row: take the other lane's
// candidate if its score is smaller, or equal with a larger truncation point.
template <int CTRL> __device__ __forceinline__ void rdoq_pick(double &j, int &l) {
    const uint64_t u = (uint64_t)__double_as_longlong(j);
    const uint32_t lo = (uint32_t)dpp_mov<CTRL>((int)(uint32_t)u), hi = (uint32_t)dpp_mov<CTRL>((int)(uint32_t)(u >> 32));
    const int ol = dpp_mov<CTRL>(l);
    const double oj = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    if (oj < j || (oj == j && ol > l)) { j = oj; l = ol; }
}

// Code region r with wavefront r.  KIND 0: A vs B; 1: A without RDOQ, any-flag;
// 2: A with RDOQ, no flag.  cb/pb/rb: LDS bases of the region's source,
// prediction and reconstruction (4-byte aligned, strides multiples of 8).
// With tab, the 8x8 basis and inverse zigzag come from that LDS copy.  LAT
// (for a caller on a serial path): the winning RDOQ lane keeps its
// reconstruction in registers instead of recomputing it, fewer instructions
// for more registers.
template <int KIND, bool LAT = false>
__device__ __forceinline__ void region_code(RegionWs &ws, int r, const uint8_t *cb, int cst,
                                            const uint8_t *pb, int pst, uint8_t *rb, int rst,
                                            int16_t *lev, uint8_t *flag, int qp, double lambda,
                                            const Tab8 *tab = nullptr) {
    const int L = threadIdx.x & 63, blk = L >> 4, c = L & 15;
    const int oy4 = (blk >> 1) * 4, ox4 = (blk & 1) * 4;
    const int m6 = qp % 6, sh = 1 << (qp / 6), qb = 15 + qp / 6, off = (1 << qb) / 3;
    const int mf0 = k_MF[m6][0], mf1 = k_MF[m6][1], mf2 = k_MF[m6][2];
    const int dv0 = k_DV[m6][0], dv1 = k_DV[m6][1], dv2 = k_DV[m6][2];
    int cu[16], pr[16], res[16], lv[16];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        const uint32_t cw_ = lds_word(cb + (oy4 + y) * cst + ox4);
        const uint32_t pw_ = lds_word(pb + (oy4 + y) * pst + ox4);
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            cu[4 * y + x] = (cw_ >> (8 * x)) & 0xff;
            pr[4 * y + x] = (pw_ >> (8 * x)) & 0xff;
            res[4 * y + x] = cu[4 * y + x] - pr[4 * y + x];
        }
    }
    {
        int co[16];
        fdct4_reg(res, co);
#pragma unroll
        for (int i = 0; i < 16; ++i) {                 // fdv.h:2430-2443
            const int mf = cls4(i) == 0 ? mf0 : cls4(i) == 1 ? mf1 : mf2;
            const int a = co[i] < 0 ? -co[i] : co[i];
            const int v = (a * mf + off) >> qb;        // |co| <= 9180: fits int32
            lv[i] = co[i] < 0 ? -v : v;
        }
    }
    int bestL = 15;
    int e4 = 0, b4 = 0, any4 = 0;
    if constexpr (LAT) {
        if (KIND != 1) {
            // lane c's own coefficient, zigzag position c: its exp-Golomb bits,
            // 8 bits per LEB byte that removing it drops, and whether it is
            // nonzero; then the block's last nonzero position, total bits and
            // drops by 16-lane DPP sums, and the drops up to c by a prefix sum
            int lvc = 0;
#pragma unroll
            for (int k = 0; k < 16; ++k) lvc = c == k ? lv[k_SCAN4c(k)] : lvc;
            const uint32_t zc = zmap(lvc);
            const int dropc = 8 * leblen(zc), gc = egbits(zc);
            const int last = row_max16(zc ? c : -1);
            const int inb = c <= last;
            const int bits0 = egbits((uint32_t)(last + 1)) + row_sum16(inb ? gc : 0);
            const int dropall = row_sum16(inb ? dropc : 0);
            const int dropto = row_prefix16(inb ? dropc : 0);
            // candidate L = c (where position c is nonzero): the levels up to c
            // kept, the bits less the drops of every position after c
            // (fdv.h:2507-2534)
            int rr[16];
            const bool valid = zc != 0;
            double jc = 0.0;
            if (valid) {
                int dq[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int dv = cls4(i) == 0 ? dv0 : cls4(i) == 1 ? dv1 : dv2;
                    dq[i] = ((k_INV4c(i) <= c ? lv[i] : 0) * dv) * sh;
                }
                idct4_reg(dq, rr);
                int D = 0;
#pragma unroll
                for (int i = 0; i < 16; ++i) { const int d = res[i] - rr[i]; D += d * d; }
                jc = (double)D + lambda * (double)(bits0 - dropall + dropto);
            }
            // L = -1: every level removed, the inverse transform of zeros is zero
            int Dm = 0;
#pragma unroll
            for (int i = 0; i < 16; ++i) Dm += res[i] * res[i];
            const double Jm = (double)Dm + lambda * (double)(bits0 - dropall);
            // The reference scans L = last down to -1, taking a score when the
            // best so far is unset or negative, or smaller.  A negative score is
            // therefore always taken and replaced by the next candidate: the
            // result is the smallest score (ties to the larger L) among the
            // candidates after the last negative one in scan order, the
            // candidate L = -1 last of all; -1 itself when its score is negative.
            int lneg = row_min16(valid && jc < 0.0 ? c : 16);
            if (Jm < 0.0) lneg = -1;
            double bj = valid && c < lneg ? jc : 1e300;
            int bl = c;
            rdoq_pick<DPP_XOR1>(bj, bl);
            rdoq_pick<DPP_XOR2>(bj, bl);
            rdoq_pick<DPP_ROR4>(bj, bl);
            rdoq_pick<DPP_ROR8>(bj, bl);
            bestL = lneg < 0 || Jm < bj ? -1 : bl;
            // the block's bits at that truncation (true exp-Golomb, not the
            // scan's), then the winner writes it: lane c == bestL, lane 0 when
            // nothing is kept
            const int bsel = egbits((uint32_t)(bestL + 1)) + row_sum16(c <= bestL ? gc : 0);
            if (c == (bestL < 0 ? 0 : bestL)) {
                uint32_t rw[4] = {0, 0, 0, 0};
                int lf[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    lf[i] = k_INV4c(i) <= bestL ? lv[i] : 0;
                    const int v = clip8(pr[i] + (bestL < 0 ? 0 : rr[i]));
                    rw[i >> 2] |= (uint32_t)v << (8 * (i & 3));
                    const int d = cu[i] - v;
                    e4 += d * d;
                }
#pragma unroll
                for (int y = 0; y < 4; ++y) *(uint32_t *)&ws.recA[r][(oy4 + y) * 8 + ox4] = rw[y];
                b4 = bsel;
                any4 = bestL >= 0;
                uint32_t *lw = (uint32_t *)&ws.levA[r][blk * 16];
#pragma unroll
                for (int i = 0; i < 8; ++i) lw[i] = ((uint32_t)(uint16_t)lf[2 * i]) | ((uint32_t)(uint16_t)lf[2 * i + 1] << 16);
            }
        } else if (c == 0) {
            int dq[16], rr[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int dv = cls4(i) == 0 ? dv0 : cls4(i) == 1 ? dv1 : dv2;
                dq[i] = (lv[i] * dv) * sh;
            }
            idct4_reg(dq, rr);
            uint32_t rw[4] = {0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int v = clip8(pr[i] + rr[i]);
                rw[i >> 2] |= (uint32_t)v << (8 * (i & 3));
                const int d = cu[i] - v;
                e4 += d * d;
            }
#pragma unroll
            for (int y = 0; y < 4; ++y) *(uint32_t *)&ws.recA[r][(oy4 + y) * 8 + ox4] = rw[y];
            int lastF = -1;
#pragma unroll
            for (int k = 0; k < 16; ++k) if (lv[k_SCAN4c(k)]) lastF = k;
            b4 = egbits((uint32_t)(lastF + 1));
#pragma unroll
            for (int k = 0; k < 16; ++k) if (k <= lastF) b4 += egbits(zmap(lv[k_SCAN4c(k)]));
            any4 = lastF >= 0;
            uint32_t *lw = (uint32_t *)&ws.levA[r][blk * 16];
#pragma unroll
            for (int i = 0; i < 8; ++i) lw[i] = ((uint32_t)(uint16_t)lv[2 * i]) | ((uint32_t)(uint16_t)lv[2 * i + 1] << 16);
        }
    } else {
        if (KIND != 1) {
            int last = -1, nzm = 0, drop[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint32_t z = zmap(lv[k_SCAN4c(k)]);
                drop[k] = 8 * leblen(z);
                if (z) { last = k; nzm |= 1 << k; }
            }
            int bits0 = egbits((uint32_t)(last + 1));
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k <= last) bits0 += egbits(zmap(lv[k_SCAN4c(k)]));
            for (int Lc = c - 1; Lc <= last; Lc += 16) {
                int bits = bits0;
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    if (k > Lc && k <= last) bits -= drop[k];
                int dq[16], rr[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int dv = cls4(i) == 0 ? dv0 : cls4(i) == 1 ? dv1 : dv2;
                    dq[i] = ((k_INV4c(i) <= Lc ? lv[i] : 0) * dv) * sh;
                }
                idct4_reg(dq, rr);
                int D = 0;
#pragma unroll
                for (int i = 0; i < 16; ++i) { const int d = res[i] - rr[i]; D += d * d; }
                ws.rdJ[r][blk][Lc + 1] = (double)D + lambda * (double)bits;
            }
            wave_sync();
            if (c == 0) {
                // the reference's scan, L = last down to -1, over scores held in registers
                double Jv[17];
#pragma unroll
                for (int k = 0; k < 17; ++k) Jv[k] = ws.rdJ[r][blk][k];
                double bestJ = -1.0;
                bestL = last;
#pragma unroll
                for (int k = 16; k >= 0; --k) {
                    const int Lc = k - 1;
                    const bool valid = Lc <= last && (Lc < 0 || ((nzm >> Lc) & 1));
                    if (valid && (bestJ < 0.0 || Jv[k] < bestJ)) { bestJ = Jv[k]; bestL = Lc; }
                }
            }
        }
        if (c == 0) {
            int lf[16], dq[16], rr[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                lf[i] = (KIND == 1 || k_INV4c(i) <= bestL) ? lv[i] : 0;
                const int dv = cls4(i) == 0 ? dv0 : cls4(i) == 1 ? dv1 : dv2;
                dq[i] = (lf[i] * dv) * sh;
            }
            idct4_reg(dq, rr);
            uint32_t rw[4] = {0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int v = clip8(pr[i] + rr[i]);
                rw[i >> 2] |= (uint32_t)v << (8 * (i & 3));
                const int d = cu[i] - v;
                e4 += d * d;
            }
#pragma unroll
            for (int y = 0; y < 4; ++y) *(uint32_t *)&ws.recA[r][(oy4 + y) * 8 + ox4] = rw[y];
            int lastF = -1;
#pragma unroll
            for (int k = 0; k < 16; ++k) if (lf[k_SCAN4c(k)]) lastF = k;
            b4 = egbits((uint32_t)(lastF + 1));
#pragma unroll
            for (int k = 0; k < 16; ++k) if (k <= lastF) b4 += egbits(zmap(lf[k_SCAN4c(k)]));
            any4 = lastF >= 0;
            uint32_t *lw = (uint32_t *)&ws.levA[r][blk * 16];
#pragma unroll
            for (int i = 0; i < 8; ++i) lw[i] = ((uint32_t)(uint16_t)lf[2 * i]) | ((uint32_t)(uint16_t)lf[2 * i + 1] << 16);
        }
    }
    // the lane of each block that holds its e4 / b4 / any4
    const bool wl = LAT && KIND != 1;
    const int w0 = wl ? max(__builtin_amdgcn_readlane(bestL, 0), 0) : 0;
    const int w1 = 16 + (wl ? max(__builtin_amdgcn_readlane(bestL, 16), 0) : 0);
    const int w2 = 32 + (wl ? max(__builtin_amdgcn_readlane(bestL, 32), 0) : 0);
    const int w3 = 48 + (wl ? max(__builtin_amdgcn_readlane(bestL, 48), 0) : 0);
    const int eA = __builtin_amdgcn_readlane(e4, w0) + __builtin_amdgcn_readlane(e4, w1) +
                   __builtin_amdgcn_readlane(e4, w2) + __builtin_amdgcn_readlane(e4, w3);
    int bA = 2 + __builtin_amdgcn_readlane(b4, w0) + __builtin_amdgcn_readlane(b4, w1) +
             __builtin_amdgcn_readlane(b4, w2) + __builtin_amdgcn_readlane(b4, w3);
    const int anyA = __builtin_amdgcn_readlane(any4, w0) | __builtin_amdgcn_readlane(any4, w1) |
                     __builtin_amdgcn_readlane(any4, w2) | __builtin_amdgcn_readlane(any4, w3);
    wave_sync();
    const int y8 = L >> 3, x8 = L & 7;
    auto T8 = [&](int i) { return tab ? tab->t8[i] : c_T8[i]; };
    if (KIND == 0) {
        const int c8 = cb[y8 * cst + x8], p8 = pb[y8 * pst + x8];
        int lv8, rr8;
        if constexpr (LAT) {
            // the transposes as lane permutes (no LDS round trips or barriers)
            const int x0 = c8 - p8;
            int a = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) a += T8(y8 * 8 + k) * __shfl(x0, k * 8 + x8);
            int Y = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) Y += __shfl(a, y8 * 8 + k) * T8(x8 * 8 + k);
            const int q = c_QS8[qp], f = q / 3;
            const int aY = Y < 0 ? -Y : Y;
            lv8 = (aY + f) / q;
            lv8 = Y < 0 ? -lv8 : lv8;
            const long long dqv = (long long)lv8 * q;
            const int dq8 = dqv > 2147483647LL ? 2147483647 : dqv < -2147483648LL ? (-2147483647 - 1) : (int)dqv;
            long long B = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) B += (long long)T8(k * 8 + y8) * __shfl(dq8, k * 8 + x8);
            const int blo = (int)(uint32_t)(uint64_t)B, bhi = (int)(uint32_t)((uint64_t)B >> 32);
            long long X = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const long long v = (long long)(((uint64_t)(uint32_t)__shfl(bhi, y8 * 8 + k) << 32) |
                                                (uint32_t)__shfl(blo, y8 * 8 + k));
                X += v * (long long)T8(k * 8 + x8);
            }
            rr8 = (int16_t)((X + (1LL << 35)) >> 36);
        } else {
            ws.x32[r][L] = c8 - p8;
            wave_sync();
            int a = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) a += T8(y8 * 8 + k) * ws.x32[r][k * 8 + x8];
            wave_sync();
            ws.x32[r][L] = a;
            wave_sync();
            int Y = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) Y += ws.x32[r][y8 * 8 + k] * T8(x8 * 8 + k);
            const int q = c_QS8[qp], f = q / 3;
            const int aY = Y < 0 ? -Y : Y;
            lv8 = (aY + f) / q;
            lv8 = Y < 0 ? -lv8 : lv8;
            const long long dqv = (long long)lv8 * q;
            const int dq8 = dqv > 2147483647LL ? 2147483647 : dqv < -2147483648LL ? (-2147483647 - 1) : (int)dqv;
            wave_sync();
            ws.x32[r][L] = dq8;
            wave_sync();
            long long B = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) B += (long long)T8(k * 8 + y8) * ws.x32[r][k * 8 + x8];
            ws.x64[r][L] = B;
            wave_sync();
            long long X = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) X += ws.x64[r][y8 * 8 + k] * (long long)T8(k * 8 + x8);
            rr8 = (int16_t)((X + (1LL << 35)) >> 36);
        }
        const int rec8 = clip8(p8 + rr8);
        const int d8 = c8 - rec8;
        const int eB = wave_sum(d8 * d8);
        const int z8 = tab ? tab->inv8[L] : k_INV8[L];
        const int last8 = wave_max(lv8 ? z8 : -1);
        int bB = 2 + egbits((uint32_t)(last8 + 1)) + wave_sum(z8 <= last8 ? egbits(zmap(lv8)) : 0);
        if (!anyA) bA = 2;
        if (last8 < 0) bB = 2;
        const bool takeA = (double)eA + lambda * (double)bA <= (double)eB + lambda * (double)bB;
        rb[y8 * rst + x8] = takeA ? ws.recA[r][L] : (uint8_t)rec8;
        lev[r * 64 + L] = takeA ? ws.levA[r][L] : (int16_t)lv8;
        if (L == 0) {
            flag[r] = takeA ? (anyA ? 1 : 0) : (last8 >= 0 ? 2 : 0);
            ws.e[r] = takeA ? eA : eB;
            ws.b[r] = takeA ? bA : bB;
        }
    } else {
        rb[y8 * rst + x8] = ws.recA[r][L];
        lev[r * 64 + L] = ws.levA[r][L];
        if (L == 0) { flag[r] = anyA ? 1 : 0; ws.e[r] = eA; ws.b[r] = bA; }
    }
}

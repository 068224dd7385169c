// fdvgpu_analyse.hip -- P-frame encode analysis and reconstruction on gfx950.
//
// One kernel, k_analyse, makes every macroblock decision of a batch of frames
// (one frame per sequence) and reconstructs them, bit-exact with the reference
// encoder's pframe_encode MB loop (yhbsh/fdv include/fdv.h:5808-6162).
//
// Parallel structure.  The MB decisions form a wavefront: an MB needs its left
// neighbour (running vector predictor, left recon column) and the MBs above and
// above-right (merge candidates, top recon row).  A 256-thread workgroup owns one
// MB row at a time and walks it left to right; rows are taken from a global
// ticket in (row, sequence) order, so a row only ever waits on rows whose
// workgroups are already running.  The row above publishes, per MB, six tagged
// 64-bit words (bottom recon rows, motion record, mode) with relaxed device-scope
// atomics; the row below spins on them.  No cache flushes are needed: a
// device-scope load sees either last frame's word (old tag) or this frame's.
//
// Inside an MB the four wavefronts work on the four 8x8 residual regions in
// parallel: 16 lanes per 4x4 block (quantisation + RDOQ) and then 64 lanes per
// 8x8 transform (int32 forward, int64 inverse), with cross-lane shuffles.
// Motion search evaluates the 8 large-diamond points at once (32 lanes each)
// from a 64x64 reference window staged in LDS.
//
// All rate-distortion costs are IEEE doubles computed without contraction
// (-ffp-contract=off) in the reference's operation order.
#include <hip/hip_runtime.h>
#include "fdvgpu_common.h"
#include "fdvgpu_dev.h"

__constant__ int32_t c_T8[64];      // 8x8 DCT basis (built on the host from libm)
__constant__ int32_t c_QS8[52];     // 8x8 quantiser steps

struct __align__(16) Smem {
    uint8_t win[2][WIN * WIN];
    uint8_t cwin[2][2][CWIN * CWIN];
    uint8_t cur[256];
    uint8_t ccur[128];                // U 8x8 then V 8x8
    uint8_t pred[256];
    uint8_t cpred[128];
    uint8_t recS[256];
    uint8_t recT[256], recI[256], recN[256], rec8[256];
    uint8_t crec[128];
    int16_t levT[256], levI[256], levN[256], lev8[256];
    int16_t clev[128];
    uint8_t flagT[4], flagI[4], flagN[4], flag8[4];
    uint8_t cflag[4];
    int32_t x32[4][64];
    int64_t x64[4][64];
    uint8_t recA[4][64], recB[4][64];
    int32_t re[4], rb[4];
    int32_t sad[2][8];
    int32_t red[2][4][4];
    int32_t satd[4];
    uint64_t hand[3][FDVG_HWORDS];
    uint8_t hbytes[48];
    uint8_t lcol[32];                 // previous MB's right recon column: Y 16, U 8, V 8
    int32_t row;
};

// ---- 4x4 transforms across a 16-lane group: lane i holds raster element i.
//      Forward core rows then columns (fdv.h:2233-2273); inverse with the >>1
//      half weights and (x+32)>>6 (fdv.h:2246-2292).
__device__ __forceinline__ int fdct4_lane(int v, int gb, int i) {
    int r = i >> 2, c = i & 3;
    int a = __shfl(v, gb + 4 * r), b = __shfl(v, gb + 4 * r + 1);
    int cc = __shfl(v, gb + 4 * r + 2), d = __shfl(v, gb + 4 * r + 3);
    int s03 = a + d, d03 = a - d, s12 = b + cc, d12 = b - cc;
    int t = c == 0 ? s03 + s12 : c == 1 ? 2 * d03 + d12 : c == 2 ? s03 - s12 : d03 - 2 * d12;
    a = __shfl(t, gb + c); b = __shfl(t, gb + 4 + c); cc = __shfl(t, gb + 8 + c); d = __shfl(t, gb + 12 + c);
    s03 = a + d; d03 = a - d; s12 = b + cc; d12 = b - cc;
    return r == 0 ? s03 + s12 : r == 1 ? 2 * d03 + d12 : r == 2 ? s03 - s12 : d03 - 2 * d12;
}
__device__ __forceinline__ int idct4_lane(int v, int gb, int i) {
    int r = i >> 2, c = i & 3;
    int a = __shfl(v, gb + 4 * r), b = __shfl(v, gb + 4 * r + 1);
    int cc = __shfl(v, gb + 4 * r + 2), d = __shfl(v, gb + 4 * r + 3);
    int e0 = a + cc, e1 = a - cc, e2 = (b >> 1) - d, e3 = b + (d >> 1);
    int t = c == 0 ? e0 + e3 : c == 1 ? e1 + e2 : c == 2 ? e1 - e2 : e0 - e3;
    a = __shfl(t, gb + c); b = __shfl(t, gb + 4 + c); cc = __shfl(t, gb + 8 + c); d = __shfl(t, gb + 12 + c);
    e0 = a + cc; e1 = a - cc; e2 = (b >> 1) - d; e3 = b + (d >> 1);
    int o = r == 0 ? e0 + e3 : r == 1 ? e1 + e2 : r == 2 ? e1 - e2 : e0 - e3;
    return (o + 32) >> 6;
}
__device__ __forceinline__ int quant4(int co, int i, int qp) {          // fdv.h:2430
    int qb = 15 + qp / 6;
    long long off = (1 << qb) / 3;
    long long a = co < 0 ? -(long long)co : co;
    int v = (int)((a * k_MF[qp % 6][pclass(i)] + off) >> qb);
    return co < 0 ? -v : v;
}
__device__ __forceinline__ int dqscale(int i, int qp) { return k_DV[qp % 6][pclass(i)]; }

// RDOQ of one 4x4 block held by 16 lanes (fdv.h:2493-2536): truncation points
// L = last..-1 scored in order, each only where L is -1 or scan position L is
// nonzero; J = SSD(res, recon of the truncated block) + lambda*bits(L), with
// bits(L) the running total the reference keeps (exp-Golomb of count and kept
// levels, minus 8 bits per LEB byte of every dropped position).  Kept when the
// best so far is negative or J is strictly smaller.
__device__ __forceinline__ int rdoq4_lane(int lv, int res, int gb, int i, int qp, double lambda) {
    const int zi = k_INV4[i];
    const int last = gmax<16>(lv ? zi : -1);
    const int mask = gor<16>(lv ? (1 << zi) : 0);
    int bits = egbits((uint32_t)(last + 1)) + gsum<16>(zi <= last ? egbits(zmap(lv)) : 0);
    const int drop = 8 * leblen(zmap(lv));
    const int Lmax = gmax<64>(last);
    const int dv = dqscale(i, qp), sh = 1 << (qp / 6);
    double bestJ = -1.0;
    int bestL = last;
    for (int L = Lmax; L >= -1; --L) {
        int tr = zi <= L ? lv : 0;
        int rr = idct4_lane((tr * dv) * sh, gb, i);
        int d = res - rr;
        int D = gsum<16>(d * d);
        bool active = L <= last;
        if (active && (L < 0 || ((mask >> L) & 1))) {
            double J = (double)D + lambda * (double)bits;
            if (bestJ < 0.0 || J < bestJ) { bestJ = J; bestL = L; }
        }
        int dL = __shfl(drop, gb + k_SCAN4[L < 0 ? 0 : L]);
        if (active && L >= 0) bits -= dL;
    }
    return zi <= bestL ? lv : 0;
}

// ---- residual regions.  Region r (one 8x8: of a luma MB, or a whole chroma
// plane block) is coded by wavefront r.  kind 0: four RDOQ'd 4x4 against one
// 8x8, ties to 4x4, empty option costs only the flag (fdv.h:5319-5410);
// kind 1: four 4x4 without RDOQ behind an any-coefficient flag (chroma of
// INTER16, fdv.h:5576-5626); kind 2: four RDOQ'd 4x4, no flag (4x4 chroma of
// INTER8, code_residual n==4, fdv.h:5306-5311).
// Pixel operands live in LDS at base + off[r] with a common stride; results:
// recon at recbase + off[r] (stride rst), levels at lev[r*64..], flag[r],
// s.re[r] (SSD), s.rb[r] (bits).  Every thread of the block must call it.
__device__ __forceinline__ void code_regions(Smem &s, int kind, int nreg,
        const uint8_t *cb, int cst, const uint8_t *pb, int pst, uint8_t *rb_, int rst,
        int o0, int o1, int o2, int o3, int16_t *lev, uint8_t *flag, int qp, double lambda) {
    const int tid = threadIdx.x, r = tid >> 6, L = tid & 63;
    const bool live = r < nreg;
    const int ro = live ? (r == 0 ? o0 : r == 1 ? o1 : r == 2 ? o2 : o3) : o0;
    const int gb = L & ~15, blk = L >> 4, i = L & 15;
    // 4x4 path
    const int y4 = (blk >> 1) * 4 + (i >> 2), x4 = (blk & 1) * 4 + (i & 3);
    const int c4 = cb[ro + y4 * cst + x4], p4 = pb[ro + y4 * pst + x4];
    const int res4 = c4 - p4;
    int lv = quant4(fdct4_lane(res4, gb, i), i, qp);
    if (kind != 1) lv = rdoq4_lane(lv, res4, gb, i, qp, lambda);
    const int rr4 = idct4_lane((lv * dqscale(i, qp)) * (1 << (qp / 6)), gb, i);
    const int rec4 = clip8(p4 + rr4);
    const int d4 = c4 - rec4;
    const int zi = k_INV4[i];
    const int last4 = gmax<16>(lv ? zi : -1);
    const int eA = gsum<64>(d4 * d4);
    int bA = 2 + gsum<64>((zi <= last4 ? egbits(zmap(lv)) : 0) + (i == 0 ? egbits((uint32_t)(last4 + 1)) : 0));
    const int anyA = gmax<64>(lv != 0);
    s.recA[r][y4 * 8 + x4] = (uint8_t)rec4;
    if (kind == 0) {
        // 8x8 path: int32 forward, quant by division, saturating dequant, int64
        // inverse (fdv.h:2614-2685)
        const int y8 = L >> 3, x8 = L & 7;
        const int c8 = cb[ro + y8 * cst + x8], p8 = pb[ro + y8 * pst + x8];
        s.x32[r][L] = c8 - p8;
        __syncthreads();
        int a = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) a += c_T8[y8 * 8 + k] * s.x32[r][k * 8 + x8];
        __syncthreads();
        s.x32[r][L] = a;
        __syncthreads();
        int Y = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) Y += s.x32[r][y8 * 8 + k] * c_T8[x8 * 8 + k];
        const int q = c_QS8[qp], f = q / 3;
        const int aY = Y < 0 ? -Y : Y;
        int lv8 = (aY + f) / q;
        lv8 = Y < 0 ? -lv8 : lv8;
        long long dqv = (long long)lv8 * q;
        const int dq8 = dqv > 2147483647LL ? 2147483647 : dqv < -2147483648LL ? (-2147483647 - 1) : (int)dqv;
        __syncthreads();
        s.x32[r][L] = dq8;
        __syncthreads();
        long long B = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) B += (long long)c_T8[k * 8 + y8] * s.x32[r][k * 8 + x8];
        s.x64[r][L] = B;
        __syncthreads();
        long long X = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) X += s.x64[r][y8 * 8 + k] * (long long)c_T8[k * 8 + x8];
        const int rr8 = (int16_t)((X + (1LL << 35)) >> 36);
        const int rec8 = clip8(p8 + rr8);
        const int d8 = c8 - rec8;
        const int eB = gsum<64>(d8 * d8);
        const int z8 = k_INV8[L];
        const int last8 = gmax<64>(lv8 ? z8 : -1);
        int bB = 2 + egbits((uint32_t)(last8 + 1)) + gsum<64>(z8 <= last8 ? egbits(zmap(lv8)) : 0);
        if (!anyA) bA = 2;
        if (last8 < 0) bB = 2;
        const bool takeA = (double)eA + lambda * (double)bA <= (double)eB + lambda * (double)bB;
        s.recB[r][L] = (uint8_t)rec8;
        __syncthreads();
        if (live) {
            rb_[ro + y8 * rst + x8] = takeA ? s.recA[r][L] : s.recB[r][L];
            if (takeA) lev[r * 64 + blk * 16 + i] = (int16_t)lv;
            else       lev[r * 64 + L] = (int16_t)lv8;
            if (L == 0) {
                flag[r] = takeA ? (anyA ? 1 : 0) : (last8 >= 0 ? 2 : 0);
                s.re[r] = takeA ? eA : eB;
                s.rb[r] = takeA ? bA : bB;
            }
        }
    } else {
        __syncthreads();
        if (live) {
            rb_[ro + (L >> 3) * rst + (L & 7)] = s.recA[r][L];
            lev[r * 64 + blk * 16 + i] = (int16_t)lv;
            if (L == 0) { flag[r] = anyA ? 1 : 0; s.re[r] = eA; s.rb[r] = bA; }
        }
    }
    __syncthreads();
}

// Block-wide sums of up to four ints; every thread gets the totals.
__device__ __forceinline__ void bsum4(Smem &s, int &par, int a, int b, int c, int d) {
    const int w = threadIdx.x >> 6, L = threadIdx.x & 63;
    a = gsum<64>(a); b = gsum<64>(b); c = gsum<64>(c); d = gsum<64>(d);
    if (L == 0) { s.red[par][w][0] = a; s.red[par][w][1] = b; s.red[par][w][2] = c; s.red[par][w][3] = d; }
    __syncthreads();
}

// ---- motion search on an LDS window.  Points are evaluated 8 at a time, 32
//      lanes each; the caller applies them in order.
struct Pts { int x[8], y[8]; int n; };

__device__ __forceinline__ void eval_pts(Smem &s, int &par, const uint8_t *w, const Pts &P,
                                         int bs, int ox, int oy) {
    const int t = threadIdx.x, k = t >> 5, tt = t & 31;
    int acc = 0;
    int mx = 0, my = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) if (j == k) { mx = P.x[j]; my = P.y[j]; }
    if (k < P.n) {
        if ((mx | my) & 3) {
            const int npx = bs * bs / 32;
            for (int j = 0; j < npx; ++j) {
                int p = tt * npx + j, y = p / bs, x = p % bs;
                int v = lumaq(w, ox + x, oy + y, mx, my);
                int d = s.cur[(oy + y) * 16 + ox + x] - v;
                acc += d < 0 ? -d : d;
            }
        } else {
            const int npx = bs * bs / 32;
            const int dx = mx >> 2, dy = my >> 2;
            for (int j = 0; j < npx; ++j) {
                int p = tt * npx + j, y = p / bs, x = p % bs;
                int v = w[(WOFF + oy + y + dy) * WIN + WOFF + ox + x + dx];
                int d = s.cur[(oy + y) * 16 + ox + x] - v;
                acc += d < 0 ? -d : d;
            }
        }
    }
    acc = gsum<32>(acc);
    if (tt == 0 && k < 8) s.sad[par][k] = acc;
    __syncthreads();
}

// EPZS diamond + sub-pel cross (fdv.h:3375-3466).  Returns the SAD; (*mx,*my)
// the quarter-pel vector.  Uniform across the block.
__device__ int me_diamond(Smem &s, int &par, const uint8_t *w, int bs, int ox, int oy,
                          int pmx, int pmy, int *omx, int *omy) {
    const int lim = FDVG_LIM;
    int px = (pmx >> 2) * 4, py = (pmy >> 2) * 4;
    px = px < -lim ? -lim : px > lim ? lim : px;
    py = py < -lim ? -lim : py > lim ? lim : py;
    Pts P;
    P.n = 2; P.x[0] = 0; P.y[0] = 0; P.x[1] = px; P.y[1] = py;
    eval_pts(s, par, w, P, bs, ox, oy);
    int best = s.sad[par][0], bxv = 0, byv = 0;
    if ((px || py) && s.sad[par][1] < best) { best = s.sad[par][1]; bxv = px; byv = py; }
    par ^= 1;
    for (int it = 0; it < 64; ++it) {
        const int cx = bxv, cy = byv;
        int valid = 0;
        P.n = 8;
        for (int k = 0; k < 8; ++k) {
            int nx = cx + k_LDSP[k][0], ny = cy + k_LDSP[k][1];
            bool ok = !(nx < -lim || nx > lim || ny < -lim || ny > lim);
            valid |= ok << k;
            P.x[k] = ok ? nx : 0; P.y[k] = ok ? ny : 0;
        }
        eval_pts(s, par, w, P, bs, ox, oy);
        int moved = 0;
        for (int k = 0; k < 8; ++k)
            if (((valid >> k) & 1) && s.sad[par][k] < best) { best = s.sad[par][k]; bxv = P.x[k]; byv = P.y[k]; moved = 1; }
        par ^= 1;
        if (!moved) break;
    }
    {
        const int cx = bxv, cy = byv;
        int valid = 0;
        P.n = 4;
        for (int k = 0; k < 4; ++k) {
            int nx = cx + k_SDSP[k][0], ny = cy + k_SDSP[k][1];
            bool ok = !(nx < -lim || nx > lim || ny < -lim || ny > lim);
            valid |= ok << k;
            P.x[k] = ok ? nx : 0; P.y[k] = ok ? ny : 0;
        }
        eval_pts(s, par, w, P, bs, ox, oy);
        for (int k = 0; k < 4; ++k)
            if (((valid >> k) & 1) && s.sad[par][k] < best) { best = s.sad[par][k]; bxv = P.x[k]; byv = P.y[k]; }
        par ^= 1;
    }
    for (int st = 2; st >= 1 && best > 0; st >>= 1) {
        const int cx = bxv, cy = byv;
        P.n = 4;
        for (int k = 0; k < 4; ++k) { P.x[k] = cx + k_CROSS[k][0] * st; P.y[k] = cy + k_CROSS[k][1] * st; }
        eval_pts(s, par, w, P, bs, ox, oy);
        for (int k = 0; k < 4; ++k)
            if (s.sad[par][k] < best) { best = s.sad[par][k]; bxv = P.x[k]; byv = P.y[k]; }
        par ^= 1;
    }
    *omx = bxv; *omy = byv;
    return best;
}

// ---- intra 16x16 / 8x8 prediction (fdv.h:2851-2911).  top[-1] is the corner.
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

// ---- handoff words
__device__ __forceinline__ uint64_t hload(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void hstore(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int hbyte(const Smem &s, int j, int b) {
    return (int)((s.hand[j][b / 7] >> (8 * (b % 7))) & 0xff);
}

// ------------------------------------------------------------------ kernel

extern "C" __global__ void __launch_bounds__(256)
k_analyse(const fdvg_seq *__restrict__ seqs, const int2 *__restrict__ rowtab, int nrows,
          unsigned *ticket, fdvg_mbrec *__restrict__ recs, uint64_t *__restrict__ hand) {
    __shared__ Smem s;
    const int tid = threadIdx.x;
    int par = 0;
    for (;;) {
        if (tid == 0) s.row = (int)atomicAdd(ticket, 1u);
        __syncthreads();
        const int row = s.row;
        __syncthreads();
        if (row >= nrows) return;
        const int2 rt = rowtab[row];
        const fdvg_seq q = seqs[rt.x];
        const int my = rt.y, mw = q.w >> 4, cw = q.w >> 1;
        const double lambda = q.lambda;
        const int qp = q.qp;
        const uint64_t tag = (uint64_t)(q.tag & 0xff) << 56;
        int pmx = 0, pmy = 0;

        for (int mx = 0; mx < mw; ++mx) {
            const int bx = mx * 16, by = my * 16, cbx = bx >> 1, cby = by >> 1;
            const int mb = q.mb0 + my * mw + mx;
            // ---- wait for the row above (MBs mx-1 .. mx+1) and stage its handoff
            if (tid < 18) {
                const int j = tid / 6, k = tid % 6, ax = mx - 1 + j;
                uint64_t v = 0;
                if (my > 0 && ax >= 0 && ax < mw) {
                    const uint64_t *p = hand + (size_t)(mb - mw - 1 + j) * FDVG_HWORDS + k;
                    v = hload(p);
                    // bounded wait: a row that never publishes is a bug, and it must
                    // surface as an error, not as a hung device
                    for (unsigned spins = 0; (v & 0xff00000000000000ull) != tag; ++spins) {
                        if (spins > (1u << 24)) { atomicOr(ticket + 1, 1u); break; }
                        __builtin_amdgcn_s_sleep(1);
                        v = hload(p);
                    }
                }
                s.hand[j][k] = v;
            }
            // ---- stage the source MB and the reference windows
            s.cur[tid] = q.cy[(size_t)(by + (tid >> 4)) * q.w + bx + (tid & 15)];
            if (tid < 128) {
                const int pl = tid >> 6, p = tid & 63;
                const uint8_t *src = pl ? q.cv : q.cu;
                s.ccur[tid] = src[(size_t)(cby + (p >> 3)) * cw + cbx + (p & 7)];
            }
            const bool need_ref = !q.force_skip && !(q.given && q.given[mb - q.mb0].mode == 2);
            const int nwin = q.force_skip ? 1 : q.navail;
            if (need_ref || q.force_skip) {
                for (int r = 0; r < nwin; ++r) {
                    // 64 rows x 64 bytes: 16 bytes a thread, two 8-byte loads (8-aligned)
                    const int wy = tid >> 2, wx = (tid & 3) * 16;
                    const uint8_t *src = q.ry[r] + (ptrdiff_t)(by - WOFF + wy) * q.ys + bx - WOFF + wx;
                    uint64_t a = *(const uint64_t *)src, b = *(const uint64_t *)(src + 8);
                    *(uint64_t *)&s.win[r][wy * WIN + wx] = a;
                    *(uint64_t *)&s.win[r][wy * WIN + wx + 8] = b;
                    // 2 planes x 32 rows x 32 bytes: 8 bytes a thread (4-aligned)
                    const int pl = tid >> 7, t2 = tid & 127, cyy = t2 >> 2, cxx = (t2 & 3) * 8;
                    const uint8_t *cp = (pl ? q.rv[r] : q.ru[r]) + (ptrdiff_t)(cby - CWOFF + cyy) * q.cs + cbx - CWOFF + cxx;
                    const uint32_t *c32 = (const uint32_t *)cp;
                    uint32_t u0 = c32[0], u1 = c32[1];
                    *(uint32_t *)&s.cwin[r][pl][cyy * CWIN + cxx] = u0;
                    *(uint32_t *)&s.cwin[r][pl][cyy * CWIN + cxx + 4] = u1;
                }
            }
            __syncthreads();

            // ---- neighbours from the row above / the previous MB
            const int ht = my > 0, hl = mx > 0;
            // merge candidates (fdv.h:5711-5732)
            int cx_[3] = {pmx, 0, 0}, cy_[3] = {pmy, 0, 0}, cr_[3] = {0, 0, 0}, ncand = 1;
            if (ht) {
                for (int k = 0; k < 2 && ncand < 3; ++k) {
                    const int j = 1 + k;
                    if (mx + k >= mw) continue;
                    const int inter = hbyte(s, j, 37);
                    if (!inter) continue;
                    const int vx = (int16_t)(hbyte(s, j, 32) | (hbyte(s, j, 33) << 8));
                    const int vy = (int16_t)(hbyte(s, j, 34) | (hbyte(s, j, 35) << 8));
                    const int vr = hbyte(s, j, 36);
                    bool dup = false;
#pragma unroll
                    for (int c = 0; c < 3; ++c) dup |= (c < ncand && cx_[c] == vx && cy_[c] == vy && cr_[c] == vr);
                    if (!dup) {
                        if (ncand == 1) { cx_[1] = vx; cy_[1] = vy; cr_[1] = vr; }
                        else            { cx_[2] = vx; cy_[2] = vy; cr_[2] = vr; }
                        ++ncand;
                    }
                }
            }

            const fdvg_mbinfo *gv = q.given ? &q.given[mb - q.mb0] : nullptr;
            const fdvg_metab *mt = (q.me_mode == FDVG_ME_TABLE && q.metab) ? &q.metab[mb - q.mb0] : nullptr;
            int mode, bmi = 0, iref = 0, imx = 0, imy = 0, nsub = 0;
            int qmv[4][2];
            fdvg_mbrec *R = &recs[mb];

            if (q.force_skip) {
                // a held frame: SKIP with the zero vector everywhere (fdv.h:5820-5836)
                mode = 0; ncand = 1; cx_[0] = 0; cy_[0] = 0; cr_[0] = 0;
                s.recS[tid] = s.win[0][(WOFF + (tid >> 4)) * WIN + WOFF + (tid & 15)];
                __syncthreads();
            } else {
                // ---- SKIP: best merge candidate by SSD, first wins (fdv.h:5847-5866)
                const int x = tid & 15, y = tid >> 4;
                int pv[3] = {0, 0, 0}, e[3] = {0, 0, 0};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (c >= ncand || (gv && !(gv->mode == 0 && gv->merge_idx == c))) continue;
                    pv[c] = lumaq(s.win[cr_[c]], x, y, cx_[c], cy_[c]);
                    const int d = s.cur[tid] - pv[c];
                    e[c] = d * d;
                }
                bsum4(s, par, e[0], e[1], e[2], 0);
                int Ds = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (c >= ncand) continue;
                    int ec = s.red[par][0][c] + s.red[par][1][c] + s.red[par][2][c] + s.red[par][3][c];
                    if (gv) { if (gv->mode == 0 && gv->merge_idx == c) { Ds = ec; bmi = c; } }
                    else if (c == 0 || ec < Ds) { Ds = ec; bmi = c; }
                }
                par ^= 1;
                s.recS[tid] = (uint8_t)(bmi == 0 ? pv[0] : bmi == 1 ? pv[1] : pv[2]);
                const double Js = (double)Ds + lambda * (double)(3 + (ncand > 1 ? 3 : 0));
                double Jbest = Js;
                bool skip_wins = Js <= lambda * (double)FL_INTRA;
                if (mt && (mt->flags & 1)) skip_wins = true;
                if (gv) skip_wins = false;

                // ---- INTER16 over the references (fdv.h:5909-5943)
                double Ji = 1e30;
                int sad16 = 0;
                const bool want_i = gv ? gv->mode == 1 : !skip_wins;
                for (int r = 0; r < q.navail && want_i; ++r) {
                    if (gv && r != gv->ref) continue;
                    if (!gv && r > 0 && Jbest <= lambda * (double)FL_INTER16) break;
                    int rmx, rmy, rsad = 0;
                    if (gv) { rmx = gv->mv[0][0]; rmy = gv->mv[0][1]; }
                    else if (mt) { rmx = mt->mv16[r][0]; rmy = mt->mv16[r][1]; rsad = mt->sad16[r]; }
                    else rsad = me_diamond(s, par, s.win[r], 16, 0, 0, pmx, pmy, &rmx, &rmy);
                    s.pred[tid] = (uint8_t)lumaq(s.win[r], tid & 15, tid >> 4, rmx, rmy);
                    __syncthreads();
                    code_regions(s, 0, 4, s.cur, 16, s.pred, 16, s.recT, 16, 0, 8, 128, 136,
                                 s.levT, s.flagT, qp, lambda);
                    const int bb = 6 + egbits(zmap(rmx - pmx)) + egbits(zmap(rmy - pmy)) +
                                   s.rb[0] + s.rb[1] + s.rb[2] + s.rb[3];
                    const double dd = (double)(s.re[0] + s.re[1] + s.re[2] + s.re[3]);
                    const double J = dd + lambda * (double)bb;
                    if (J < Jbest) Jbest = J;
                    if (J < Ji) {
                        Ji = J; iref = r; imx = rmx; imy = rmy; sad16 = rsad;
                        s.recI[tid] = s.recT[tid];
                        s.levI[tid] = s.levT[tid];
                        if (tid < 4) s.flagI[tid] = s.flagT[tid];
                    }
                    __syncthreads();
                }

                // ---- INTRA 16x16: SATD screen, one coding (fdv.h:5950-5988)
                double Jn = 1e30;
                const uint8_t *top = &s.hbytes[1];
                if (tid < 16) s.hbytes[1 + tid] = ht ? (uint8_t)hbyte(s, 1, tid) : 0;
                if (tid == 16) s.hbytes[0] = (ht && hl) ? (uint8_t)hbyte(s, 0, 15) : 128;
                __syncthreads();
                const bool want_n = gv ? gv->mode == 2 : (!skip_wins && Jbest > lambda * (double)FL_INTRA);
                if (want_n) {
                    const uint8_t *tp = top, *lft = s.lcol;
                    const Nb nb = nb_setup(tp, lft, s.hbytes[0], 16, ht, hl);
                    int sel = 0;
                    if (gv) sel = gv->sub;
                    else {
                        if (tid < 64) {
                            const int m = tid >> 4, b = tid & 15, oy = (b >> 2) * 4, ox = (b & 3) * 4;
                            int dd4[16];
#pragma unroll
                            for (int k = 0; k < 16; ++k)
                                dd4[k] = s.cur[(oy + (k >> 2)) * 16 + ox + (k & 3)] -
                                         ipred(m, nb, tp, lft, 16, ox + (k & 3), oy + (k >> 2));
                            int t4[16], sa = 0;
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                int p0 = dd4[4 * k] + dd4[4 * k + 1], m0 = dd4[4 * k] - dd4[4 * k + 1];
                                int p1 = dd4[4 * k + 2] + dd4[4 * k + 3], m1 = dd4[4 * k + 2] - dd4[4 * k + 3];
                                t4[4 * k] = p0 + p1; t4[4 * k + 1] = m0 + m1; t4[4 * k + 2] = p0 - p1; t4[4 * k + 3] = m0 - m1;
                            }
#pragma unroll
                            for (int k = 0; k < 4; ++k) {
                                int p0 = t4[k] + t4[4 + k], m0 = t4[k] - t4[4 + k];
                                int p1 = t4[8 + k] + t4[12 + k], m1 = t4[8 + k] - t4[12 + k];
                                sa += abs(p0 + p1) + abs(m0 + m1) + abs(p0 - p1) + abs(m0 - m1);
                            }
                            sa = gsum<16>(sa);
                            if (b == 0) s.satd[m] = sa;
                        }
                        __syncthreads();
                        int sbest = 0; sel = -1;
                        for (int m = 0; m < 4; ++m) {
                            if (!nn_ok(m, nb)) continue;
                            if (sel < 0 || s.satd[m] < sbest) { sbest = s.satd[m]; sel = m; }
                        }
                    }
                    s.pred[tid] = (uint8_t)ipred(sel, nb, tp, lft, 16, tid & 15, tid >> 4);
                    __syncthreads();
                    code_regions(s, 0, 4, s.cur, 16, s.pred, 16, s.recN, 16, 0, 8, 128, 136,
                                 s.levN, s.flagN, qp, lambda);
                    const int bn = 6 + s.rb[0] + s.rb[1] + s.rb[2] + s.rb[3];
                    Jn = (double)(s.re[0] + s.re[1] + s.re[2] + s.re[3]) + lambda * (double)bn;
                    nsub = sel;
                    if (Jn < Jbest) Jbest = Jn;
                }

                // ---- INTER8: four chained quadrant searches, the SAD split gate,
                //      four 8x8 residuals (fdv.h:5996-6033)
                double J8 = 1e30;
                bool want_8 = gv ? gv->mode == 3 : (!skip_wins && Jbest > lambda * (double)FL_INTER8);
                if (want_8) {
                    int ssad = 0, sx = pmx, sy = pmy;
                    for (int qd = 0; qd < 4; ++qd) {
                        int a, b;
                        if (gv) { a = gv->mv[qd][0]; b = gv->mv[qd][1]; }
                        else if (mt) { a = mt->mv8[qd][0]; b = mt->mv8[qd][1]; ssad += mt->sad8[qd]; }
                        else ssad += me_diamond(s, par, s.win[0], 8, (qd & 1) * 8, (qd >> 1) * 8, sx, sy, &a, &b);
                        qmv[qd][0] = a; qmv[qd][1] = b; sx = a; sy = b;
                    }
                    if (!gv && sad16 > 0 && ssad * 10 >= sad16 * 9) want_8 = false;
                    if (want_8) {
                        const int x = tid & 15, y = tid >> 4, qd = (y >> 3) * 2 + (x >> 3);
                        const int vx = qd == 0 ? qmv[0][0] : qd == 1 ? qmv[1][0] : qd == 2 ? qmv[2][0] : qmv[3][0];
                        const int vy = qd == 0 ? qmv[0][1] : qd == 1 ? qmv[1][1] : qd == 2 ? qmv[2][1] : qmv[3][1];
                        s.pred[tid] = (uint8_t)lumaq(s.win[0], x, y, vx, vy);
                        __syncthreads();
                        code_regions(s, 0, 4, s.cur, 16, s.pred, 16, s.rec8, 16, 0, 8, 128, 136,
                                     s.lev8, s.flag8, qp, lambda);
                        int b8 = 3 + s.rb[0] + s.rb[1] + s.rb[2] + s.rb[3];
                        int qpx = pmx, qpy = pmy;
                        for (int qd2 = 0; qd2 < 4; ++qd2) {
                            b8 += egbits(zmap(qmv[qd2][0] - qpx)) + egbits(zmap(qmv[qd2][1] - qpy));
                            qpx = qmv[qd2][0]; qpy = qmv[qd2][1];
                        }
                        J8 = (double)(s.re[0] + s.re[1] + s.re[2] + s.re[3]) + lambda * (double)b8;
                    }
                }

                // ---- commit: strict < in order SKIP, INTER16, INTRA, INTER8
                if (gv) mode = gv->mode;
                else {
                    double Jm = Js; mode = 0;
                    if (Ji < Jm) { Jm = Ji; mode = 1; }
                    if (Jn < Jm) { Jm = Jn; mode = 2; }
                    if (J8 < Jm) { Jm = J8; mode = 3; }
                }
            }

            const int selx = bmi == 0 ? cx_[0] : bmi == 1 ? cx_[1] : cx_[2];
            const int sely = bmi == 0 ? cy_[0] : bmi == 1 ? cy_[1] : cy_[2];
            const int selr = bmi == 0 ? cr_[0] : bmi == 1 ? cr_[1] : cr_[2];
            // ---- chroma of the chosen mode, then publish the MB
            uint8_t *yrec = mode == 0 ? s.recS : mode == 1 ? s.recI : mode == 2 ? s.recN : s.rec8;
            int ckind = 0;
            if (mode == 0) {
                const int sr = selr, smx = selx / 2, smy = sely / 2;
                if (tid < 128) {
                    const int pl = tid >> 6, p = tid & 63;
                    s.crec[tid] = (uint8_t)chromaq(s.cwin[sr][pl], p & 7, p >> 3, smx, smy);
                }
                __syncthreads();
            } else if (mode == 1) {
                if (tid < 128) {
                    const int pl = tid >> 6, p = tid & 63;
                    s.cpred[tid] = (uint8_t)chromaq(s.cwin[iref][pl], p & 7, p >> 3, imx / 2, imy / 2);
                }
                __syncthreads();
                code_regions(s, 1, 2, s.ccur, 8, s.cpred, 8, s.crec, 8, 0, 64, 0, 0, s.clev, s.cflag, qp, lambda);
                ckind = 1;
            } else if (mode == 2) {
                if (tid < 2) {
                    // chroma neighbours: top from the handoff, left from the previous MB
                    const int pl = tid;
                    uint8_t *ct = &s.hbytes[18 + pl * 10];
                    for (int k = 0; k < 8; ++k) ct[1 + k] = ht ? (uint8_t)hbyte(s, 1, 16 + 8 * pl + k) : 0;
                    ct[0] = (ht && hl) ? (uint8_t)hbyte(s, 0, 16 + 8 * pl + 7) : 128;
                }
                __syncthreads();
                if (tid < 128) {
                    const int pl = tid >> 6, p = tid & 63;
                    const uint8_t *ct = &s.hbytes[18 + pl * 10];
                    const uint8_t *tp = ct + 1, *lf = &s.lcol[16 + 8 * pl];
                    const Nb nb = nb_setup(tp, lf, ct[0], 8, ht, hl);
                    s.cpred[tid] = (uint8_t)ipred(nsub, nb, tp, lf, 8, p & 7, p >> 3);
                }
                __syncthreads();
                code_regions(s, 0, 2, s.ccur, 8, s.cpred, 8, s.crec, 8, 0, 64, 0, 0, s.clev, s.cflag, qp, lambda);
                ckind = 2;
            } else {
                if (tid < 128) {
                    const int pl = tid >> 6, p = tid & 63, x = p & 7, y = p >> 3, qd = (y >> 2) * 2 + (x >> 2);
                    const int vx = qd == 0 ? qmv[0][0] : qd == 1 ? qmv[1][0] : qd == 2 ? qmv[2][0] : qmv[3][0];
                    const int vy = qd == 0 ? qmv[0][1] : qd == 1 ? qmv[1][1] : qd == 2 ? qmv[2][1] : qmv[3][1];
                    s.cpred[tid] = (uint8_t)chromaq(s.cwin[0][pl], x, y, vx / 2, vy / 2);
                }
                __syncthreads();
                code_regions(s, 2, 2, s.ccur, 8, s.cpred, 8, s.crec, 8, 0, 64, 0, 0, s.clev, s.cflag, qp, lambda);
                ckind = 3;
            }

            // reconstruction out (the next frame's reference, deblocked later)
            q.oy[(ptrdiff_t)(by + (tid >> 4)) * q.ys + bx + (tid & 15)] = yrec[tid];
            if (tid < 128) {
                const int pl = tid >> 6, p = tid & 63;
                (pl ? q.ov : q.ou)[(ptrdiff_t)(cby + (p >> 3)) * q.cs + cbx + (p & 7)] = s.crec[tid];
            }
            // decision record for the serialiser
            const int16_t *lsrc = mode == 1 ? s.levI : mode == 2 ? s.levN : s.lev8;
            const uint8_t *fsrc = mode == 1 ? s.flagI : mode == 2 ? s.flagN : s.flag8;
            R->lev[tid] = mode == 0 ? (int16_t)0 : lsrc[tid];
            if (tid < 128) R->clev[tid] = ckind ? s.clev[tid] : (int16_t)0;
            // new vector predictor and motion record
            int nmx, nmy, nref, ninter;
            if (mode == 0)      { nmx = selx; nmy = sely; nref = selr; ninter = 1; }
            else if (mode == 1) { nmx = imx; nmy = imy; nref = iref; ninter = 1; }
            else if (mode == 2) { nmx = 0; nmy = 0; nref = 0; ninter = 0; }
            else                { nmx = qmv[3][0]; nmy = qmv[3][1]; nref = 0; ninter = 1; }
            if (tid == 0) {
                R->mode = (uint8_t)mode; R->ref = (uint8_t)(mode == 0 ? selr : mode == 1 ? iref : 0);
                R->merge_idx = (uint8_t)bmi; R->ncand = (uint8_t)ncand;
                R->sub = (uint8_t)nsub; R->ckind = (uint8_t)ckind;
                for (int k = 0; k < 4; ++k) R->lflag[k] = mode == 0 ? 0 : fsrc[k];
                R->cflag[0] = s.cflag[0]; R->cflag[1] = s.cflag[1];
                for (int k = 0; k < 4; ++k) { R->mvd[k][0] = R->mvd[k][1] = 0; R->mv[k][0] = R->mv[k][1] = 0; }
                if (mode == 0) { R->mv[0][0] = (int16_t)nmx; R->mv[0][1] = (int16_t)nmy; }
                if (mode == 1) {
                    R->mv[0][0] = (int16_t)imx; R->mv[0][1] = (int16_t)imy;
                    R->mvd[0][0] = (int16_t)(imx - pmx); R->mvd[0][1] = (int16_t)(imy - pmy);
                }
                if (mode == 3) {
                    int qpx = pmx, qpy = pmy;
                    for (int k = 0; k < 4; ++k) {
                        R->mv[k][0] = (int16_t)qmv[k][0]; R->mv[k][1] = (int16_t)qmv[k][1];
                        R->mvd[k][0] = (int16_t)(qmv[k][0] - qpx); R->mvd[k][1] = (int16_t)(qmv[k][1] - qpy);
                        qpx = qmv[k][0]; qpy = qmv[k][1];
                    }
                }
                if (q.info) {
                    fdvg_mbinfo *I = &q.info[mb - q.mb0];
                    I->mode = (uint8_t)mode; I->ref = R->ref; I->merge_idx = (uint8_t)bmi; I->sub = (uint8_t)nsub;
                    for (int k = 0; k < 4; ++k) { I->mv[k][0] = R->mv[k][0]; I->mv[k][1] = R->mv[k][1]; }
                }
            }
            // handoff bytes: bottom rows, motion record, mode
            if (tid < 16) s.hbytes[tid] = yrec[240 + tid];
            else if (tid < 24) s.hbytes[tid] = s.crec[56 + (tid - 16)];
            else if (tid < 32) s.hbytes[tid] = s.crec[64 + 56 + (tid - 24)];
            else if (tid == 32) {
                s.hbytes[32] = (uint8_t)(nmx & 0xff); s.hbytes[33] = (uint8_t)((nmx >> 8) & 0xff);
                s.hbytes[34] = (uint8_t)(nmy & 0xff); s.hbytes[35] = (uint8_t)((nmy >> 8) & 0xff);
                s.hbytes[36] = (uint8_t)nref; s.hbytes[37] = (uint8_t)ninter; s.hbytes[38] = (uint8_t)mode;
                for (int k = 39; k < 42; ++k) s.hbytes[k] = 0;
            }
            // left column for the next MB
            if (tid >= 64 && tid < 80) s.lcol[tid - 64] = yrec[(tid - 64) * 16 + 15];
            else if (tid >= 80 && tid < 88) s.lcol[16 + tid - 80] = s.crec[(tid - 80) * 8 + 7];
            else if (tid >= 88 && tid < 96) s.lcol[24 + tid - 88] = s.crec[64 + (tid - 88) * 8 + 7];
            __syncthreads();
            if (tid < FDVG_HWORDS) {
                uint64_t v = tag;
                for (int b = 0; b < 7; ++b) v |= (uint64_t)s.hbytes[tid * 7 + b] << (8 * b);
                hstore(hand + (size_t)mb * FDVG_HWORDS + tid, v);
            }
            if (mode != 2) { pmx = nmx; pmy = nmy; }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------ host launchers

extern "C" int fdvg_set_tables(const int32_t *c8, const int32_t *qs8) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_T8), c8, 64 * sizeof(int32_t)) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_QS8), qs8, 52 * sizeof(int32_t)) != hipSuccess) return -1;
    return 0;
}

// Workgroups to launch: enough to keep every row in flight that can be, capped
// by what the device holds at once (the ticket loop takes the rest).
extern "C" int fdvg_analyse_grid(int nrows) {
    static int cap = 0;
    if (!cap) {
        int dev = 0, ncu = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess) return -1;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_analyse, 256, 0) != hipSuccess) return -1;
        cap = ncu * (per > 0 ? per : 1);
    }
    return nrows < cap ? nrows : cap;
}

extern "C" int fdvg_launch_analyse(const fdvg_seq *seqs, const int2 *rowtab, int nrows,
                                   unsigned *ticket, fdvg_mbrec *recs, uint64_t *hand,
                                   hipStream_t st) {
    int grid = fdvg_analyse_grid(nrows);
    if (grid <= 0) return -1;
    hipLaunchKernelGGL(k_analyse, dim3(grid), dim3(256), 0, st, seqs, rowtab, nrows, ticket, recs, hand);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

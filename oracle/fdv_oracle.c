/* fdv_oracle.c -- CPU restatement of fdv's P-frame encode-analysis path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU-baseline port); the
 * product never links it.  See fdv_oracle.h for the contract.  Citations are
 * to yhbsh/fdv include/fdv.h.
 *
 * Written as a restatement of the algorithm, not of the reference's code
 * layout: motion compensation is evaluated per sample from the H.264
 * definitions, RDOQ is stated as a scan over independently scored truncation
 * points, and the P-frame driver keeps each candidate as levels + region flags
 * and serialises only the winner.  Results must be byte-identical to the
 * reference; the parity tests pin that against oracle/_ref and fixtures.
 *
 * Build with -ffp-contract=off: rate-distortion costs are IEEE doubles and a
 * fused multiply-add would change tie-breaks.
 */
#include "fdv_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ bits -- */

static int clip8(int v) { return v < 0 ? 0 : v > 255 ? 255 : v; }
static int iabs(int v) { return v < 0 ? -v : v; }

/* Signed -> unsigned zigzag (fdv.h:1701). */
static uint32_t zmap(int v) { return ((uint32_t)v << 1) ^ (uint32_t)(v >> 31); }

/* LEB128 (fdv.h:1704-1708). */
static int leb_len(uint32_t u) { int n = 1; while (u >= 0x80) { u >>= 7; ++n; } return n; }
static size_t leb_write(uint8_t *b, size_t p, uint32_t u) {
    while (u >= 0x80) { b[p++] = (uint8_t)(0x80 | (u & 0x7f)); u >>= 7; }
    b[p++] = (uint8_t)u;
    return p;
}

/* Exp-Golomb length 2*floor(log2(v+1))+1 (fdv_bits_val, fdv.h:2487). */
int orc_bits_val(uint32_t v) {
    uint64_t x = (uint64_t)v + 1;
    int lg = 0;
    while (x >> (lg + 1)) ++lg;
    return 2 * lg + 1;
}

/* --------------------------------------------------------------- tables -- */

/* Scan position -> raster index (fdv.h:1699 and 2153). */
static const uint8_t SCAN4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
static const uint8_t SCAN8[64] = {
     0,  1,  8, 16,  9,  2,  3, 10, 17, 24, 32, 25, 18, 11,  4,  5,
    12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,  6,  7, 14, 21, 28,
    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
    58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

/* H.264 forward multipliers and dequant scales by [qp%6][position class]
 * (fdv.h:2405-2421): class 0 = both coordinates even, 1 = both odd, 2 = mixed. */
static const int QMF[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490}, {10082, 4194, 6554},
                              {9362, 3647, 5825},  {8192, 3355, 5243},  {7282, 2893, 4559}};
static const int QDV[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16},
                              {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
static int pclass(int i) {
    int r = (i >> 2) & 1, c = i & 1;
    return (r == c) ? r : 2;          /* even/even -> 0, odd/odd -> 1 (fdv.h:2423) */
}

/* 8x8 DCT basis and step tables, built from libm exactly as c8_init
 * (fdv.h:2597-2612): basis a_u cos((2x+1)u pi/16) scaled by 512 and rounded,
 * AC rows then made zero-sum by adjusting column 0; step 262144*0.625*2^(qp/6). */
static int32_t T8[8][8];
static int32_t QS8[52];
static pthread_once_t t8_once = PTHREAD_ONCE_INIT;
static void t8_build(void) {
    const double pi = 3.14159265358979323846;
    for (int u = 0; u < 8; ++u) {
        double a = u ? sqrt(2.0 / 8.0) : sqrt(1.0 / 8.0);
        for (int x = 0; x < 8; ++x)
            T8[u][x] = (int32_t)lround(a * cos((2 * x + 1) * u * pi / 16.0) * 512.0);
    }
    for (int u = 1; u < 8; ++u) {
        int32_t s = 0;
        for (int x = 0; x < 8; ++x) s += T8[u][x];
        T8[u][0] -= s;
    }
    for (int q = 0; q < 52; ++q)
        QS8[q] = (int32_t)lround(262144.0 * (0.625 * pow(2.0, q / 6.0)));
}
void orc_tables8(int32_t c8[64], int32_t qstep8[52]) {
    pthread_once(&t8_once, t8_build);
    memcpy(c8, T8, sizeof T8);
    memcpy(qstep8, QS8, sizeof QS8);
}

/* ------------------------------------------------------- 4x4 transforms -- */

/* Forward core: rows then columns of [1 1 1 1; 2 1 -1 -2; 1 -1 -1 1; 1 -2 2 -1]
 * (fwd_1d / fdv_fdct4x4_scalar, fdv.h:2233, 2257). */
static void f4(int a, int b, int c, int d, int o[4]) {
    int s03 = a + d, d03 = a - d, s12 = b + c, d12 = b - c;
    o[0] = s03 + s12;
    o[1] = 2 * d03 + d12;
    o[2] = s03 - s12;
    o[3] = d03 - 2 * d12;
}
void orc_fdct4(const int16_t r[16], int32_t c[16]) {
    int t[16], o[4];
    for (int i = 0; i < 4; ++i) {
        f4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3], o);
        for (int j = 0; j < 4; ++j) t[4 * i + j] = o[j];
    }
    for (int j = 0; j < 4; ++j) {
        f4(t[j], t[4 + j], t[8 + j], t[12 + j], o);
        for (int i = 0; i < 4; ++i) c[4 * i + j] = o[i];
    }
}

/* Inverse core with the >>1 half-weights, rows then columns, then (x+32)>>6
 * (inv_1d / fdv_idct4x4_scalar, fdv.h:2246, 2275). */
static void g4(int a, int b, int c, int d, int o[4]) {
    int e0 = a + c, e1 = a - c, e2 = (b >> 1) - d, e3 = b + (d >> 1);
    o[0] = e0 + e3; o[1] = e1 + e2; o[2] = e1 - e2; o[3] = e0 - e3;
}
void orc_idct4(const int32_t c[16], int16_t r[16]) {
    int t[16], o[4];
    for (int i = 0; i < 4; ++i) {
        g4(c[4 * i], c[4 * i + 1], c[4 * i + 2], c[4 * i + 3], o);
        for (int j = 0; j < 4; ++j) t[4 * i + j] = o[j];
    }
    for (int j = 0; j < 4; ++j) {
        g4(t[j], t[4 + j], t[8 + j], t[12 + j], o);
        for (int i = 0; i < 4; ++i) r[4 * i + j] = (int16_t)((o[i] + 32) >> 6);
    }
}

/* Dead-zone quantiser, rounding offset a third of a step (fdv.h:2430-2443). */
void orc_quant4(const int32_t c[16], int16_t l[16], int qp) {
    int qb = 15 + qp / 6;
    int64_t off = (1 << qb) / 3;
    for (int i = 0; i < 16; ++i) {
        int64_t a = c[i] < 0 ? -(int64_t)c[i] : c[i];
        int v = (int)((a * QMF[qp % 6][pclass(i)] + off) >> qb);
        l[i] = (int16_t)(c[i] < 0 ? -v : v);
    }
}

/* level * V * 2^(qp/6) in int32 (fdv_dequant4x4_scalar, fdv.h:2546). */
void orc_dequant4(const int16_t l[16], int32_t c[16], int qp) {
    for (int i = 0; i < 16; ++i)
        c[i] = (int32_t)(l[i] * QDV[qp % 6][pclass(i)]) * (int32_t)(1 << (qp / 6));
}

/* RD truncation of the trailing scan positions (fdv_rdoq4x4_inner,
 * fdv.h:2493-2536).  Candidates are "keep scan positions 0..L" for L = last
 * down to -1, scored only where L is -1 or position L holds a nonzero level.
 * Score J(L) = SSD(res, recon of the truncated block) + lambda * bits(L), where
 * bits(last) is the exp-Golomb length of the count plus that of every kept
 * level, and each step down subtracts 8 bits per LEB byte of the dropped
 * position (zeros included), exactly as the reference's running total does.
 * The scan keeps a candidate when the best so far is negative or J is strictly
 * smaller -- the reference's `bestJ < 0 || J < bestJ` with bestJ starting at
 * -1, which also means a negative score is always replaced by the next one. */
void orc_rdoq4(int16_t l[16], const int16_t res[16], int qp, double lambda) {
    int last = -1;
    for (int k = 0; k < 16; ++k) if (l[SCAN4[k]]) last = k;
    if (last < 0) return;
    int bits = orc_bits_val((uint32_t)(last + 1));
    for (int k = 0; k <= last; ++k) bits += orc_bits_val(zmap(l[SCAN4[k]]));
    double bestJ = -1.0;
    int bestL = last;
    for (int L = last; L >= -1; --L) {
        if (L < 0 || l[SCAN4[L]]) {
            int16_t tr[16], rr[16];
            int32_t dq[16];
            for (int i = 0; i < 16; ++i) tr[i] = 0;
            for (int k = 0; k <= L; ++k) tr[SCAN4[k]] = l[SCAN4[k]];
            orc_dequant4(tr, dq, qp);
            orc_idct4(dq, rr);
            int64_t D = 0;
            for (int i = 0; i < 16; ++i) { int64_t d = res[i] - rr[i]; D += d * d; }
            double J = (double)D + lambda * (double)bits;
            if (bestJ < 0.0 || J < bestJ) { bestJ = J; bestL = L; }
        }
        if (L >= 0) bits -= 8 * leb_len(zmap(l[SCAN4[L]]));
    }
    for (int k = bestL + 1; k < 16; ++k) l[SCAN4[k]] = 0;
}

/* ------------------------------------------------------- 8x8 transforms -- */

/* Y = T X T^T in int32, columns first (fdv_fdct8x8_inner, fdv.h:2614). */
void orc_fdct8(const int16_t x[64], int32_t y[64]) {
    pthread_once(&t8_once, t8_build);
    int32_t a[64];
    for (int u = 0; u < 8; ++u)
        for (int c = 0; c < 8; ++c) {
            int32_t s = 0;
            for (int r = 0; r < 8; ++r) s += T8[u][r] * (int32_t)x[8 * r + c];
            a[8 * u + c] = s;
        }
    for (int u = 0; u < 8; ++u)
        for (int v = 0; v < 8; ++v) {
            int32_t s = 0;
            for (int c = 0; c < 8; ++c) s += a[8 * u + c] * T8[v][c];
            y[8 * u + v] = s;
        }
}

/* X = (T^T Y T + 2^35) >> 36 with int64 accumulation (fdv.h:2639-2654). */
void orc_idct8(const int32_t y[64], int16_t x[64]) {
    pthread_once(&t8_once, t8_build);
    int64_t b[64];
    for (int r = 0; r < 8; ++r)
        for (int v = 0; v < 8; ++v) {
            int64_t s = 0;
            for (int u = 0; u < 8; ++u) s += (int64_t)T8[u][r] * y[8 * u + v];
            b[8 * r + v] = s;
        }
    for (int r = 0; r < 8; ++r)
        for (int c = 0; c < 8; ++c) {
            int64_t s = 0;
            for (int v = 0; v < 8; ++v) s += b[8 * r + v] * T8[v][c];
            x[8 * r + c] = (int16_t)((s + ((int64_t)1 << 35)) >> 36);
        }
}

/* sign * ((|c| + q/3) / q) with integer division (fdv.h:2664-2673). */
void orc_quant8(const int32_t y[64], int16_t l[64], int qp) {
    pthread_once(&t8_once, t8_build);
    int32_t q = QS8[qp], f = q / 3;
    for (int i = 0; i < 64; ++i) {
        int32_t a = y[i] < 0 ? -y[i] : y[i];
        int32_t v = (a + f) / q;
        l[i] = (int16_t)(y[i] < 0 ? -v : v);
    }
}

/* level * q saturated to int32 (fdv.h:2675-2685). */
void orc_dequant8(const int16_t l[64], int32_t y[64], int qp) {
    pthread_once(&t8_once, t8_build);
    for (int i = 0; i < 64; ++i) {
        int64_t v = (int64_t)l[i] * QS8[qp];
        y[i] = v > INT32_MAX ? INT32_MAX : v < INT32_MIN ? INT32_MIN : (int32_t)v;
    }
}

/* 4x4 Hadamard SATD of src - pred (fdv_satd4x4_scalar, fdv.h:3803). */
int orc_satd4(const uint8_t *src, int ss, const uint8_t *pred) {
    int d[16], t[16];
    for (int i = 0; i < 16; ++i) d[i] = src[(i >> 2) * ss + (i & 3)] - pred[i];
    for (int i = 0; i < 4; ++i) {
        int p0 = d[4 * i] + d[4 * i + 1], m0 = d[4 * i] - d[4 * i + 1];
        int p1 = d[4 * i + 2] + d[4 * i + 3], m1 = d[4 * i + 2] - d[4 * i + 3];
        t[4 * i] = p0 + p1; t[4 * i + 1] = m0 + m1; t[4 * i + 2] = p0 - p1; t[4 * i + 3] = m0 - m1;
    }
    int s = 0;
    for (int j = 0; j < 4; ++j) {
        int p0 = t[j] + t[4 + j], m0 = t[j] - t[4 + j];
        int p1 = t[8 + j] + t[12 + j], m1 = t[8 + j] - t[12 + j];
        s += iabs(p0 + p1) + iabs(m0 + m1) + iabs(p0 - p1) + iabs(m0 - m1);
    }
    return s;
}

/* ----------------------------------------------------------------- intra -- */

static int nn_ok(int m, int n, int ht, int hl) {          /* fdv.h:2827 */
    if (m == 0) return 1;
    if (m == 1) return ht;
    if (m == 2) return hl;
    if (m == 3) return ht && hl && (n == 8 || n == 16 || n == 32);
    return 0;
}

/* DC / V / H / PLANE prediction of an n x n block (fdv.h:2851-2911). */
void orc_intra_nxn(int mode, const uint8_t *top, const uint8_t *left, uint8_t tl, int n,
                   int ht, int hl, uint8_t *pred) {
    if (mode == 3 && nn_ok(3, n, ht, hl)) {
        int hn = n / 2, H = 0, V = 0;
        for (int k = 1; k <= hn; ++k) {
            int i = hn - 1 - k;
            H += k * (top[hn - 1 + k] - (i < 0 ? tl : top[i]));
            V += k * (left[hn - 1 + k] - (i < 0 ? tl : left[i]));
        }
        int a = 16 * (left[n - 1] + top[n - 1]), b, c;
        if (n == 32)      { b = (11 * H + 512) >> 10; c = (11 * V + 512) >> 10; }
        else if (n == 16) { b = (5 * H + 32) >> 6;    c = (5 * V + 32) >> 6; }
        else              { b = (17 * H + 16) >> 5;   c = (17 * V + 16) >> 5; }
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x)
                pred[y * n + x] = (uint8_t)clip8((a + c * (y - (hn - 1)) + 16 + b * (x - (hn - 1))) >> 5);
        return;
    }
    if (mode == 1 && ht) {
        for (int y = 0; y < n; ++y) memcpy(pred + y * n, top, (size_t)n);
        return;
    }
    if (mode == 2 && hl) {
        for (int y = 0; y < n; ++y) memset(pred + y * n, left[y], (size_t)n);
        return;
    }
    int dc = 128, s = 0;
    if (ht && hl) { for (int k = 0; k < n; ++k) s += top[k] + left[k]; dc = (s + n) / (2 * n); }
    else if (ht)  { for (int k = 0; k < n; ++k) s += top[k];           dc = (s + n / 2) / n; }
    else if (hl)  { for (int k = 0; k < n; ++k) s += left[k];          dc = (s + n / 2) / n; }
    memset(pred, dc, (size_t)n * n);
}

/* ----------------------------------------------------------------- inter -- */

/* Sample-wise H.264 quarter-pel luma interpolation (fdv.h:2931-2990, the
 * per-sample form the block path fdv_mc_luma_inner 3082 is equivalent to).
 * Names follow H.264 figure 8-4: b/h horizontal/vertical half samples,
 * j the centre, G integer; quarter samples average two neighbours. */
#define PX(x, y) ((int)ref[(ptrdiff_t)(y) * stride + (x)])
static int tap6(int a, int b, int c, int d, int e, int f) {
    return a - 5 * b + 20 * c + 20 * d - 5 * e + f;
}
static int hraw(const uint8_t *ref, int stride, int x, int y) {
    return tap6(PX(x - 2, y), PX(x - 1, y), PX(x, y), PX(x + 1, y), PX(x + 2, y), PX(x + 3, y));
}
static int vraw(const uint8_t *ref, int stride, int x, int y) {
    return tap6(PX(x, y - 2), PX(x, y - 1), PX(x, y), PX(x, y + 1), PX(x, y + 2), PX(x, y + 3));
}
static int sb(const uint8_t *ref, int stride, int x, int y) { return clip8((hraw(ref, stride, x, y) + 16) >> 5); }
static int sh(const uint8_t *ref, int stride, int x, int y) { return clip8((vraw(ref, stride, x, y) + 16) >> 5); }
static int sj(const uint8_t *ref, int stride, int x, int y) {
    int v = tap6(hraw(ref, stride, x, y - 2), hraw(ref, stride, x, y - 1), hraw(ref, stride, x, y),
                 hraw(ref, stride, x, y + 1), hraw(ref, stride, x, y + 2), hraw(ref, stride, x, y + 3));
    return clip8((v + 512) >> 10);
}
static int avg2(int a, int b) { return (a + b + 1) >> 1; }
static int luma_sample(const uint8_t *ref, int stride, int x, int y, int fx, int fy) {
    switch (4 * fy + fx) {
    case 0:  return PX(x, y);
    case 1:  return avg2(PX(x, y), sb(ref, stride, x, y));
    case 2:  return sb(ref, stride, x, y);
    case 3:  return avg2(sb(ref, stride, x, y), PX(x + 1, y));
    case 4:  return avg2(PX(x, y), sh(ref, stride, x, y));
    case 5:  return avg2(sb(ref, stride, x, y), sh(ref, stride, x, y));
    case 6:  return avg2(sb(ref, stride, x, y), sj(ref, stride, x, y));
    case 7:  return avg2(sb(ref, stride, x, y), sh(ref, stride, x + 1, y));
    case 8:  return sh(ref, stride, x, y);
    case 9:  return avg2(sh(ref, stride, x, y), sj(ref, stride, x, y));
    case 10: return sj(ref, stride, x, y);
    case 11: return avg2(sj(ref, stride, x, y), sh(ref, stride, x + 1, y));
    case 12: return avg2(sh(ref, stride, x, y), PX(x, y + 1));
    case 13: return avg2(sh(ref, stride, x, y), sb(ref, stride, x, y + 1));
    case 14: return avg2(sj(ref, stride, x, y), sb(ref, stride, x, y + 1));
    default: return avg2(sh(ref, stride, x + 1, y), sb(ref, stride, x, y + 1));
    }
}

void orc_mc_luma(const uint8_t *ref, int stride, int bx, int by, int bw, int bh,
                 int mvx, int mvy, uint8_t *dst, int dstride) {
    int ix = bx + (mvx >> 2), iy = by + (mvy >> 2), fx = mvx & 3, fy = mvy & 3;
    for (int i = 0; i < bh; ++i)
        for (int j = 0; j < bw; ++j)
            dst[i * dstride + j] = (uint8_t)luma_sample(ref, stride, ix + j, iy + i, fx, fy);
}

/* Bilinear eighth-weight chroma interpolation at quarter-sample phase
 * (fdv_mc_chroma_inner, fdv.h:3260-3283). */
void orc_mc_chroma(const uint8_t *ref, int stride, int bx, int by, int bw, int bh,
                   int mvx, int mvy, uint8_t *dst, int dstride) {
    int ix = bx + (mvx >> 2), iy = by + (mvy >> 2), fx = mvx & 3, fy = mvy & 3;
    for (int i = 0; i < bh; ++i)
        for (int j = 0; j < bw; ++j) {
            int x = ix + j, y = iy + i;
            int top = PX(x, y) * (4 - fx) + PX(x + 1, y) * fx;
            int bot = PX(x, y + 1) * (4 - fx) + PX(x + 1, y + 1) * fx;
            dst[i * dstride + j] = (uint8_t)((top * (4 - fy) + bot * fy + 8) >> 4);
        }
}
#undef PX

/* SAD of the bw x bh block of cur at (bx,by) against the reference displaced
 * by a quarter-pel vector (block_sad, fdv.h:3321). */
static int sad_at(const uint8_t *cur, int cs, const uint8_t *ref, int rs,
                  int bx, int by, int bw, int bh, int mvx, int mvy) {
    uint8_t p[256];
    orc_mc_luma(ref, rs, bx, by, bw, bh, mvx, mvy, p, bw);
    int s = 0;
    for (int i = 0; i < bh; ++i)
        for (int j = 0; j < bw; ++j) s += iabs(cur[(by + i) * cs + bx + j] - p[i * bw + j]);
    return s;
}

/* Sub-pel refinement: half- then quarter-pel four-point cross around the best
 * point, unclamped, stopping at a perfect match (fdv.h:3453-3461). */
static void subpel_refine(const uint8_t *cur, int cs, const uint8_t *ref, int rs,
                          int bx, int by, int bw, int bh, int *bx_, int *by_, int *best) {
    static const int8_t X4[4][2] = {{0, -1}, {0, 1}, {-1, 0}, {1, 0}};
    for (int st = 2; st >= 1 && *best > 0; st >>= 1) {
        int cx = *bx_, cy = *by_;
        for (int k = 0; k < 4; ++k) {
            int nx = cx + X4[k][0] * st, ny = cy + X4[k][1] * st;
            int s = sad_at(cur, cs, ref, rs, bx, by, bw, bh, nx, ny);
            if (s < *best) { *best = s; *bx_ = nx; *by_ = ny; }
        }
    }
}

/* Motion search, quarter-pel vectors (fdv_me_search_inner, fdv.h:3375-3466).
 * Diamond: start at (0,0), try the predictor rounded down to whole pels and
 * clamped to +-range, then up to 64 large-diamond rounds (8 points, strict
 * improvement, first wins) and one small-diamond round, all inside +-range;
 * then the sub-pel cross.  Exhaustive: every whole-pel point in +-range, ties
 * to the smaller |dx|+|dy| and then raster order, then the same sub-pel cross. */
int orc_me_search(const uint8_t *cur, int cs, const uint8_t *ref, int rs,
                  int bx, int by, int bw, int bh, int range, int pmx, int pmy,
                  int me_mode, int *mvx, int *mvy) {
    int lim = range * 4;
    int bxv = 0, byv = 0, best;
    if (me_mode == ORC_ME_EXHAUSTIVE) {
        best = -1;
        int bn = 0;
        for (int dy = -lim; dy <= lim; dy += 4)
            for (int dx = -lim; dx <= lim; dx += 4) {
                int s = sad_at(cur, cs, ref, rs, bx, by, bw, bh, dx, dy);
                int nrm = iabs(dx) + iabs(dy);
                if (best < 0 || s < best || (s == best && nrm < bn)) {
                    best = s; bn = nrm; bxv = dx; byv = dy;
                }
            }
    } else {
        static const int8_t BIG[8][2] = {{0, 8}, {0, -8}, {8, 0}, {-8, 0},
                                         {4, 4}, {4, -4}, {-4, 4}, {-4, -4}};
        static const int8_t SMALL[4][2] = {{0, 4}, {0, -4}, {4, 0}, {-4, 0}};
        best = sad_at(cur, cs, ref, rs, bx, by, bw, bh, 0, 0);
        int px = (pmx >> 2) * 4, py = (pmy >> 2) * 4;
        px = px < -lim ? -lim : px > lim ? lim : px;
        py = py < -lim ? -lim : py > lim ? lim : py;
        if (px || py) {
            int s = sad_at(cur, cs, ref, rs, bx, by, bw, bh, px, py);
            if (s < best) { best = s; bxv = px; byv = py; }
        }
        for (int it = 0; it < 64; ++it) {
            int cx = bxv, cy = byv, moved = 0;
            for (int k = 0; k < 8; ++k) {
                int nx = cx + BIG[k][0], ny = cy + BIG[k][1];
                if (nx < -lim || nx > lim || ny < -lim || ny > lim) continue;
                int s = sad_at(cur, cs, ref, rs, bx, by, bw, bh, nx, ny);
                if (s < best) { best = s; bxv = nx; byv = ny; moved = 1; }
            }
            if (!moved) break;
        }
        int cx = bxv, cy = byv;
        for (int k = 0; k < 4; ++k) {
            int nx = cx + SMALL[k][0], ny = cy + SMALL[k][1];
            if (nx < -lim || nx > lim || ny < -lim || ny > lim) continue;
            int s = sad_at(cur, cs, ref, rs, bx, by, bw, bh, nx, ny);
            if (s < best) { best = s; bxv = nx; byv = ny; }
        }
    }
    subpel_refine(cur, cs, ref, rs, bx, by, bw, bh, &bxv, &byv, &best);
    *mvx = bxv; *mvy = byv;
    return best;
}

/* --------------------------------------------------------------- deblock -- */

/* One edge: p1 p0 | q0 q1 (filter_edge, fdv.h:3491-3503). */
static void edge(uint8_t *q, ptrdiff_t st, int al, int be, int tc) {
    int p1 = q[-2 * st], p0 = q[-st], q0 = q[0], q1 = q[st];
    if (iabs(p0 - q0) >= al || iabs(p1 - p0) >= be || iabs(q1 - q0) >= be) return;
    int d = ((q0 - p0) * 4 + (p1 - q1) + 4) >> 3;
    d = d < -tc ? -tc : d > tc ? tc : d;
    q[-st] = (uint8_t)clip8(p0 + d);
    q[0] = (uint8_t)clip8(q0 - d);
}

/* All vertical 8x8-grid edges, then all horizontal ones, in place
 * (fdv_deblock_plane_inner scalar branch, fdv.h:3617-3625). */
void orc_deblock_plane(uint8_t *p, int w, int h, int stride, int qp) {
    int al = 6 + qp, be = 2 + qp / 4, tc = 1 + qp / 10;
    for (int y = 0; y < h; ++y)
        for (int x = 8; x < w; x += 8) edge(p + (ptrdiff_t)y * stride + x, 1, al, be, tc);
    for (int y = 8; y < h; y += 8)
        for (int x = 0; x < w; ++x) edge(p + (ptrdiff_t)y * stride + x, stride, al, be, tc);
}

/* ---------------------------------------------------------------- planes -- */

int orc_plane_stride(int w) { return (w + 2 * ORC_BORDER + 31) & ~31; }

void orc_plane_extend(uint8_t *d, int w, int h, int st) {
    for (int y = 0; y < h; ++y) {
        uint8_t *r = d + (ptrdiff_t)y * st;
        memset(r - ORC_BORDER, r[0], ORC_BORDER);
        memset(r + w, r[w - 1], ORC_BORDER);
    }
    for (int k = 1; k <= ORC_BORDER; ++k) {
        memcpy(d - (ptrdiff_t)k * st - ORC_BORDER, d - ORC_BORDER, (size_t)w + 2 * ORC_BORDER);
        memcpy(d + (ptrdiff_t)(h - 1 + k) * st - ORC_BORDER,
               d + (ptrdiff_t)(h - 1) * st - ORC_BORDER, (size_t)w + 2 * ORC_BORDER);
    }
}

void orc_plane_from_tight(uint8_t *d, int st, const uint8_t *src, int w, int h) {
    for (int y = 0; y < h; ++y) memcpy(d + (ptrdiff_t)y * st, src + (size_t)y * w, (size_t)w);
    orc_plane_extend(d, w, h, st);
}

/* ------------------------------------------------------- residual coding -- */

/* A coded n x n residual kept in symbolic form until a mode wins: one region
 * flag per aligned 8x8 (0 empty, 1 four 4x4, 2 one 8x8) and the levels of the
 * chosen transform in raster order of their blocks.  For n == 4 there is a
 * single flagless 4x4 block. */
typedef struct {
    int n, nreg;
    uint8_t flag[4];
    int16_t lev[4][64];       /* region r: 4 x 16 (flag 1) or 64 (flag 2) */
} coded_t;

/* Byte sinks for the 16 streams. */
typedef struct { uint8_t *b[ORC_NS]; size_t n[ORC_NS]; size_t cap; int over; } sink_t;
static void put_byte(sink_t *s, int k, unsigned v) {
    if (s->n[k] >= s->cap) { s->over = 1; return; }
    s->b[k][s->n[k]++] = (uint8_t)v;
}
static void put_leb(sink_t *s, int k, uint32_t u) {
    if (s->n[k] + 5 > s->cap) { s->over = 1; return; }
    s->n[k] = leb_write(s->b[k], s->n[k], u);
}
/* count + levels of one 4x4 block in scan order, to the (chroma or luma)
 * 4x4 streams (code_block4 emission, fdv.h:5268-5277). */
static void emit4(sink_t *s, const int16_t lev[16], int chroma) {
    int last = -1;
    for (int k = 0; k < 16; ++k) if (lev[SCAN4[k]]) last = k;
    put_byte(s, chroma ? ORC_S_CNC : ORC_S_CN, (unsigned)(last + 1));
    for (int k = 0; k <= last; ++k) put_leb(s, chroma ? ORC_S_CLC : ORC_S_CL, zmap(lev[SCAN4[k]]));
}
/* fdv_coeff8_encode (fdv.h:2197). */
static void emit8(sink_t *s, const int16_t lev[64]) {
    int last = -1;
    for (int k = 0; k < 64; ++k) if (lev[SCAN8[k]]) last = k;
    put_byte(s, ORC_S_CN8, (unsigned)(last + 1));
    for (int k = 0; k <= last; ++k) put_leb(s, ORC_S_CL8, zmap(lev[SCAN8[k]]));
}
/* Serialise a coded residual in the reference's stream order
 * (fdv.h:5385-5410 for regions; 5306-5311 for n == 4). */
static void emit_coded(sink_t *s, const coded_t *c, int chroma) {
    if (c->n == 4) { emit4(s, c->lev[0], chroma); return; }
    for (int r = 0; r < c->nreg; ++r) {
        put_byte(s, ORC_S_FL, c->flag[r]);
        if (c->flag[r] == 1)
            for (int b = 0; b < 4; ++b) emit4(s, c->lev[r] + 16 * b, chroma);
        else if (c->flag[r] == 2)
            emit8(s, c->lev[r]);
    }
}

/* One 4x4 block through fdct, quant, RDOQ, dequant, idct: recon, SSD and the
 * block's rate (code_block4, fdv.h:5214-5278). */
static void block4(const uint8_t *cur, int cs, const uint8_t *pred, int ps, uint8_t *rec, int rs,
                   int qp, double lambda, int16_t lev[16], int32_t *ssd, int *bits) {
    int16_t res[16], rr[16];
    int32_t co[16], dq[16];
    for (int i = 0; i < 16; ++i) res[i] = (int16_t)(cur[(i >> 2) * cs + (i & 3)] - pred[(i >> 2) * ps + (i & 3)]);
    orc_fdct4(res, co);
    orc_quant4(co, lev, qp);
    orc_rdoq4(lev, res, qp, lambda);
    orc_dequant4(lev, dq, qp);
    orc_idct4(dq, rr);
    int32_t e = 0;
    for (int i = 0; i < 16; ++i) {
        int r = clip8(pred[(i >> 2) * ps + (i & 3)] + rr[i]);
        rec[(i >> 2) * rs + (i & 3)] = (uint8_t)r;
        int d = cur[(i >> 2) * cs + (i & 3)] - r;
        e += d * d;
    }
    *ssd = e;
    int last = -1;
    for (int k = 0; k < 16; ++k) if (lev[SCAN4[k]]) last = k;
    int b = orc_bits_val((uint32_t)(last + 1));
    for (int k = 0; k <= last; ++k) b += orc_bits_val(zmap(lev[SCAN4[k]]));
    *bits = b;
}

/* One aligned 8x8 region of a residual: the cheaper in D + lambda*R of four
 * RDOQ'd 4x4 transforms and one plain 8x8 transform, ties to the 4x4 option;
 * an option that codes no level costs only the flag (the loop body of
 * fdv_code_residual_inner, fdv.h:5322-5410).  cb / pb are the region's source
 * and prediction, rb gets its reconstruction; flag 0 empty, 1 four 4x4 (lev
 * holds them as four raster blocks of 16), 2 one 8x8 (lev raster). */
static void region8(const uint8_t *cb, int cs, const uint8_t *pb, int ps, int qp, double lambda,
                    uint8_t *rb, int rs, uint8_t *flag, int16_t lev[64], double *e, int *b) {
    /* option A: four 4x4 */
    uint8_t recA[64];
    int16_t levA[64];
    double eA = 0.0; int bA = 2, anyA = 0;    /* FDV_BITS_FLAG (fdv.h:3354) */
    for (int s = 0; s < 4; ++s) {
        int sx = (s & 1) * 4, sy = (s >> 1) * 4;
        int32_t e4; int b4;
        block4(cb + sy * cs + sx, cs, pb + sy * ps + sx, ps, recA + sy * 8 + sx, 8,
               qp, lambda, levA + 16 * s, &e4, &b4);
        eA += (double)e4; bA += b4;
        for (int i = 0; i < 16; ++i) anyA |= levA[16 * s + i] != 0;
    }
    /* option B: one 8x8, no RDOQ */
    int16_t res8[64], lev8[64], rr8[64];
    int32_t co8[64], dq8[64];
    for (int i = 0; i < 64; ++i) res8[i] = (int16_t)(cb[(i >> 3) * cs + (i & 7)] - pb[(i >> 3) * ps + (i & 7)]);
    orc_fdct8(res8, co8);
    orc_quant8(co8, lev8, qp);
    orc_dequant8(lev8, dq8, qp);
    orc_idct8(dq8, rr8);
    uint8_t recB[64];
    int32_t eBi = 0;
    for (int i = 0; i < 64; ++i) {
        int v = clip8(pb[(i >> 3) * ps + (i & 7)] + rr8[i]);
        recB[i] = (uint8_t)v;
        int d = cb[(i >> 3) * cs + (i & 7)] - v;
        eBi += d * d;
    }
    double eB = (double)eBi;
    int last8 = -1;
    for (int k = 0; k < 64; ++k) if (lev8[SCAN8[k]]) last8 = k;
    int bB = 2 + orc_bits_val((uint32_t)(last8 + 1));
    for (int k = 0; k <= last8; ++k) bB += orc_bits_val(zmap(lev8[SCAN8[k]]));
    if (!anyA) bA = 2;
    if (last8 < 0) bB = 2;
    const int takeA = eA + lambda * bA <= eB + lambda * bB;
    *flag = takeA ? (anyA ? 1 : 0) : (last8 >= 0 ? 2 : 0);
    memcpy(lev, takeA ? levA : lev8, 64 * sizeof *lev);
    for (int i = 0; i < 64; ++i) rb[(i >> 3) * rs + (i & 7)] = takeA ? recA[i] : recB[i];
    *e = takeA ? eA : eB;
    *b = takeA ? bA : bB;
}

/* An n x n residual (n = 16, 8 or 4), region by region in raster order
 * (fdv_code_residual_inner, fdv.h:5300-5414).  pred is n x n packed; rec gets
 * the n x n reconstruction.  Adds to *ssd and *bits like the reference. */
static void residual(const uint8_t *cur, int cs, int ox, int oy, const uint8_t *pred, int n,
                     int qp, double lambda, uint8_t *rec, coded_t *out, double *ssd, int *bits) {
    out->n = n;
    if (n == 4) {
        int32_t e; int b;
        out->nreg = 1;
        block4(cur + (ptrdiff_t)oy * cs + ox, cs, pred, 4, rec, 4, qp, lambda, out->lev[0], &e, &b);
        *ssd += (double)e; *bits += b;
        return;
    }
    double et = 0.0; int bt = 0;
    out->nreg = (n / 8) * (n / 8);
    int r = 0;
    for (int ry = 0; ry < n; ry += 8)
        for (int rx = 0; rx < n; rx += 8, ++r) {
            double e; int b;
            region8(cur + (ptrdiff_t)(oy + ry) * cs + ox + rx, cs, pred + ry * n + rx, n, qp, lambda,
                    rec + ry * n + rx, n, &out->flag[r], out->lev[r], &e, &b);
            et += e; bt += b;
        }
    *ssd += et; *bits += bt;
}

void orc_code_residual(const uint8_t *cur, int cstride, int ox, int oy,
                       const uint8_t *pred, int n, int qp, double lambda, int chroma,
                       uint8_t *rec, uint8_t *const bufs[7], size_t lens[7],
                       double *ssd, int *bits) {
    coded_t c;
    residual(cur, cstride, ox, oy, pred, n, qp, lambda, rec, &c, ssd, bits);
    uint8_t *sb_[ORC_NS] = {0};
    uint8_t dummy[1];
    for (int k = 0; k < ORC_NS; ++k) sb_[k] = dummy;
    sb_[ORC_S_CN] = bufs[0]; sb_[ORC_S_CL] = bufs[1]; sb_[ORC_S_CN8] = bufs[2];
    sb_[ORC_S_CL8] = bufs[3]; sb_[ORC_S_FL] = bufs[4]; sb_[ORC_S_CNC] = bufs[5];
    sb_[ORC_S_CLC] = bufs[6];
    sink_t s;
    memset(&s, 0, sizeof s);
    for (int k = 0; k < ORC_NS; ++k) s.b[k] = sb_[k];
    s.cap = (size_t)1 << 30;
    emit_coded(&s, &c, chroma);
    lens[0] = s.n[ORC_S_CN]; lens[1] = s.n[ORC_S_CL]; lens[2] = s.n[ORC_S_CN8];
    lens[3] = s.n[ORC_S_CL8]; lens[4] = s.n[ORC_S_FL]; lens[5] = s.n[ORC_S_CNC];
    lens[6] = s.n[ORC_S_CLC];
}

/* An inter chroma 8x8 block: MC, and unless SKIP four 4x4 transforms without
 * RDOQ behind one any-coefficient flag (chroma_block_inner, fdv.h:5576-5626). */
static void chroma_inter(const uint8_t *cur, int cw, uint8_t *rec, const uint8_t *ref, int rs,
                         int cbx, int cby, int mvx, int mvy, int qp, int skip, sink_t *s) {
    uint8_t pred[64];
    orc_mc_chroma(ref, rs, cbx, cby, 8, 8, mvx, mvy, pred, 8);
    if (skip) {
        for (int i = 0; i < 64; ++i) rec[(cby + (i >> 3)) * cw + cbx + (i & 7)] = pred[i];
        return;
    }
    int16_t lev[4][16];
    int any = 0;
    for (int b = 0; b < 4; ++b) {
        int sx = (b & 1) * 4, sy = (b >> 1) * 4;
        int16_t res[16], rr[16];
        int32_t co[16], dq[16];
        for (int i = 0; i < 16; ++i)
            res[i] = (int16_t)(cur[(cby + sy + (i >> 2)) * cw + cbx + sx + (i & 3)] -
                               pred[(sy + (i >> 2)) * 8 + sx + (i & 3)]);
        orc_fdct4(res, co);
        orc_quant4(co, lev[b], qp);
        orc_dequant4(lev[b], dq, qp);
        orc_idct4(dq, rr);
        for (int i = 0; i < 16; ++i)
            rec[(cby + sy + (i >> 2)) * cw + cbx + sx + (i & 3)] =
                (uint8_t)clip8(pred[(sy + (i >> 2)) * 8 + sx + (i & 3)] + rr[i]);
        for (int i = 0; i < 16; ++i) any |= lev[b][i] != 0;
    }
    put_byte(s, ORC_S_FL, any ? 1 : 0);
    if (any) for (int b = 0; b < 4; ++b) emit4(s, lev[b], 1);
}

/* Neighbours of an n x n block from a tight reconstruction (gather_nb,
 * fdv.h:5201-5212): top row, left column, corner (128 unless both exist). */
static void neighbours(const uint8_t *buf, int st, int bx, int by, int n, int ytop,
                       uint8_t *top, uint8_t *left, uint8_t *tl, int *ht, int *hl) {
    *ht = by > ytop; *hl = bx > 0;
    memset(top, 0, (size_t)n); memset(left, 0, (size_t)n);
    if (*ht) memcpy(top, buf + (ptrdiff_t)(by - 1) * st + bx, (size_t)n);
    if (*hl) for (int k = 0; k < n; ++k) left[k] = buf[(ptrdiff_t)(by + k) * st + bx - 1];
    *tl = (*ht && *hl) ? buf[(ptrdiff_t)(by - 1) * st + bx - 1] : 128;
}

/* ------------------------------------------------------------ the P-frame -- */

typedef struct { int16_t x, y; uint8_t ref, inter; } mvrec_t;   /* fdv_mbmv 5680 */

double orc_lambda(int qp, double lscale) {
    return lscale * 0.425 * pow(2.0, (qp - 12) / 3.0);             /* fdv.h:5804, 3372 */
}

/* Merge list: running predictor (ref 0), then the inter MBs above and
 * above-right, deduplicated on (x, y, ref), at most 3 (merge_cands, fdv.h:5711). */
static int merge_list(const mvrec_t *map, int mw, int mx, int my, int pmx, int pmy, mvrec_t *c) {
    int n = 0;
    c[n++] = (mvrec_t){(int16_t)pmx, (int16_t)pmy, 0, 1};
    for (int k = 0; k < 2 && n < 3; ++k) {
        int x = mx + k, y = my - 1;
        if (y < 0 || x >= mw) continue;
        const mvrec_t *m = &map[y * mw + x];
        if (!m->inter) continue;
        int dup = 0;
        for (int j = 0; j < n; ++j) dup |= (c[j].x == m->x && c[j].y == m->y && c[j].ref == m->ref);
        if (!dup) c[n++] = *m;
    }
    return n;
}

/* Test hook: with a nonzero value orc_pframe_analyse returns the reconstruction
 * before the deblocking pass (the planes the MB loop's intra neighbours read). */
static _Thread_local int orc_no_deblock = 0;
void orc_set_no_deblock(int on) { orc_no_deblock = on; }

/* The MB loop of pframe_encode.  With `mt` the motion vectors and SADs come
 * from the per-MB table instead of orc_me_search (16x16 per reference, the four
 * quadrants on reference 0), and an MB whose flags bit 0 is set weighs SKIP
 * against INTER16 only. */
static int analyse_core(const uint8_t *cy, const uint8_t *cu, const uint8_t *cv,
                        int w, int h,
                        const uint8_t *const RY[2], const uint8_t *const RU[2],
                        const uint8_t *const RV[2], int ys, int cs,
                        int navail, int qp, double lambda, int force_skip, int me_mode,
                        const orc_metab *mt, const orc_mbinfo *given,
                        uint8_t *ry, uint8_t *ru, uint8_t *rv, orc_mbinfo *info,
                        uint8_t *const bufs[ORC_NS], size_t scap, size_t sn[ORC_NS]) {
    const int cw = w / 2, mw = w / 16, mh = h / 16;
    /* Exact floors (fdv.h:5887-5896) with FDV_BITS_MODE 3, FDV_BITS_FLAG 2. */
    const int FL_INTRA = 2 * 3 + 4 * 2, FL_INTER16 = 2 * 3 + 2 + 4 * 2, FL_INTER8 = 3 + 4 * 2 + 4 * 2;
    uint8_t *mmap = calloc((size_t)mw * mh, 1);
    mvrec_t *mvm = calloc((size_t)mw * mh, sizeof *mvm);
    /* structure symbols with their model index, in coding order */
    size_t scap_s = (size_t)mw * mh * 32 + 64;
    uint8_t *st_sym = malloc(scap_s), *st_ctx = malloc(scap_s);
    if (!mmap || !mvm || !st_sym || !st_ctx) { free(mmap); free(mvm); free(st_sym); free(st_ctx); return -1; }
    size_t sp = 0;
    sink_t S;
    memset(&S, 0, sizeof S);
    for (int k = 0; k < ORC_NS; ++k) S.b[k] = bufs[k];
    S.cap = scap;
#define STRUCT(ctx, v) do { st_ctx[sp] = (uint8_t)(ctx); st_sym[sp++] = (uint8_t)(v); } while (0)

    for (int my = 0; my < mh; ++my) {
        int pmx = 0, pmy = 0;
        for (int mx = 0; mx < mw; ++mx) {
            const int bx = mx * 16, by = my * 16, cbx = bx / 2, cby = by / 2;
            const int ctx = (mx > 0 && mmap[my * mw + mx - 1] ? 2 : 0) +
                            (my > 0 && mmap[(my - 1) * mw + mx] ? 1 : 0);   /* mb_mode_ctx 4172 */
            orc_mbinfo *mi = info ? &info[my * mw + mx] : NULL;
            if (force_skip) {                                   /* fdv.h:5820-5836 */
                STRUCT(ctx, 0);
                mmap[my * mw + mx] = 0;
                mvm[my * mw + mx] = (mvrec_t){0, 0, 0, 1};
                for (int i = 0; i < 16; ++i) memcpy(ry + (by + i) * w + bx, RY[0] + (ptrdiff_t)(by + i) * ys + bx, 16);
                chroma_inter(NULL, cw, ru, RU[0], cs, cbx, cby, 0, 0, qp, 1, &S);
                chroma_inter(NULL, cw, rv, RV[0], cs, cbx, cby, 0, 0, qp, 1, &S);
                if (mi) *mi = (orc_mbinfo){0, 0, 0, 0, {{0, 0}}};
                continue;
            }
            const orc_mbinfo *gv = given ? &given[my * mw + mx] : NULL;
            const orc_metab *tm = mt && !gv ? &mt[my * mw + mx] : NULL;
            const int pruned = tm && (tm->flags & 1);      /* SKIP vs INTER16 only */

            /* SKIP: best merge candidate by SSD, first wins (fdv.h:5847-5866) */
            mvrec_t cand[3];
            int ncand = merge_list(mvm, mw, mx, my, pmx, pmy, cand);
            uint8_t predS[256];
            int bmi = 0; int32_t Ds = 0;
            for (int c = 0; c < ncand; ++c) {
                if (gv && !(gv->mode == 0 && gv->merge_idx == c)) continue;
                uint8_t tp[256];
                orc_mc_luma(RY[cand[c].ref], ys, bx, by, 16, 16, cand[c].x, cand[c].y, tp, 16);
                int32_t e = 0;
                for (int i = 0; i < 256; ++i) { int d = cy[(by + (i >> 4)) * w + bx + (i & 15)] - tp[i]; e += d * d; }
                if ((gv && gv->mode == 0) || c == 0 || e < Ds) { Ds = e; bmi = c; memcpy(predS, tp, 256); }
            }
            double Js = (double)Ds + lambda * (double)(3 + (ncand > 1 ? 3 : 0));
            double Jbest = Js;
            int skip_wins = Js <= lambda * FL_INTRA;

            /* INTER16 over the available references (fdv.h:5909-5943) */
            double Ji = 1e30; int ibest_ref = 0, imx = 0, imy = 0, sad16 = 0;
            uint8_t recI[256]; coded_t codI;
            int want_i = !skip_wins, want_n, want_8;
            if (gv) { skip_wins = 0; want_i = gv->mode == 1; }
            for (int r = 0; r < navail && want_i; ++r) {
                if (gv && r != gv->ref) continue;
                if (!gv && r > 0 && Jbest <= lambda * FL_INTER16) break;
                int rmx, rmy, rsad;
                if (gv) { rmx = gv->mv[0][0]; rmy = gv->mv[0][1]; rsad = 0; }
                else if (tm) { rmx = tm->mv16[r][0]; rmy = tm->mv16[r][1]; rsad = tm->sad16[r]; }
                else rsad = orc_me_search(cy, w, RY[r], ys, bx, by, 16, 16, 16, pmx, pmy, me_mode, &rmx, &rmy);
                uint8_t pr[256], rr[256];
                orc_mc_luma(RY[r], ys, bx, by, 16, 16, rmx, rmy, pr, 16);
                coded_t tc;
                double dd = 0.0;
                int bb = 3 + 3 + orc_bits_val(zmap(rmx - pmx)) + orc_bits_val(zmap(rmy - pmy));
                residual(cy, w, bx, by, pr, 16, qp, lambda, rr, &tc, &dd, &bb);
                double J = dd + lambda * bb;
                if (J < Jbest) Jbest = J;
                if (J < Ji) {
                    Ji = J; ibest_ref = r; imx = rmx; imy = rmy; sad16 = rsad;
                    memcpy(recI, rr, 256); codI = tc;
                }
            }

            /* INTRA: SATD screen over DC/V/H/PLANE, then one coding (fdv.h:5950-5988) */
            double Jn = 1e30; int nsub = 0;
            uint8_t recN[256]; coded_t codN;
            uint8_t ntop[16], nleft[16], ntl; int nht, nhl;
            neighbours(ry, w, bx, by, 16, 0, ntop, nleft, &ntl, &nht, &nhl);
            want_n = gv ? gv->mode == 2 : (!skip_wins && !pruned && Jbest > lambda * FL_INTRA);
            if (want_n) {
                uint8_t pI[256];
                int sel = -1, sbest = 0;
                if (gv && gv->sub != ORC_SUB_SCREEN) sel = gv->sub;
                else
                    for (int m = 0; m < 4; ++m) {
                        if (!nn_ok(m, 16, nht, nhl)) continue;
                        orc_intra_nxn(m, ntop, nleft, ntl, 16, nht, nhl, pI);
                        int sa = 0;
                        for (int i = 0; i < 16; i += 4)
                            for (int j = 0; j < 16; j += 4) {
                                uint8_t pk[16];
                                for (int k = 0; k < 4; ++k) memcpy(pk + 4 * k, pI + (i + k) * 16 + j, 4);
                                sa += orc_satd4(cy + (by + i) * w + bx + j, w, pk);
                            }
                        if (sel < 0 || sa < sbest) { sbest = sa; sel = m; }
                    }
                if (sel >= 0) {
                    orc_intra_nxn(sel, ntop, nleft, ntl, 16, nht, nhl, pI);
                    double dn = 0.0; int bn = 6;
                    residual(cy, w, bx, by, pI, 16, qp, lambda, recN, &codN, &dn, &bn);
                    Jn = dn + lambda * bn;
                    nsub = sel;
                    if (Jn < Jbest) Jbest = Jn;
                }
            }

            /* INTER8: four chained quadrant searches, SAD split gate, then four
             * 8x8 residuals (fdv.h:5996-6033) */
            double J8 = 1e30;
            int qmv[4][2];
            uint8_t rec8[256]; coded_t cod8[4];
            uint8_t s8[64]; int s8n = 0;      /* the four (dx, dy) LEB pairs */
            uint8_t s8ctx[64];
            want_8 = gv ? gv->mode == 3 : (!skip_wins && !pruned && Jbest > lambda * FL_INTER8);
            if (want_8) {
                int ssad = 0, sx = pmx, syv = pmy;
                for (int q = 0; q < 4; ++q) {
                    int qx = (q & 1) * 8, qy = (q >> 1) * 8, a, b;
                    if (gv) { a = gv->mv[q][0]; b = gv->mv[q][1]; }
                    else if (tm) { a = tm->mv8[q][0]; b = tm->mv8[q][1]; ssad += tm->sad8[q]; }
                    else ssad += orc_me_search(cy, w, RY[0], ys, bx + qx, by + qy, 8, 8, 16, sx, syv, me_mode, &a, &b);
                    qmv[q][0] = a; qmv[q][1] = b; sx = a; syv = b;
                }
                if (!gv && sad16 > 0 && ssad * 10 >= sad16 * 9) want_8 = 0;
                if (want_8) {
                    double d8 = 0.0; int b8 = 3;
                    int qpx = pmx, qpy = pmy;
                    for (int q = 0; q < 4; ++q) {
                        int qx = (q & 1) * 8, qy = (q >> 1) * 8;
                        uint32_t zx = zmap(qmv[q][0] - qpx), zy = zmap(qmv[q][1] - qpy);
                        int p0 = s8n;
                        s8n = (int)leb_write(s8, (size_t)s8n, zx);
                        for (int k = p0; k < s8n; ++k) s8ctx[k] = ORC_S_MVX;
                        p0 = s8n;
                        s8n = (int)leb_write(s8, (size_t)s8n, zy);
                        for (int k = p0; k < s8n; ++k) s8ctx[k] = ORC_S_MVY;
                        b8 += orc_bits_val(zx) + orc_bits_val(zy);
                        qpx = qmv[q][0]; qpy = qmv[q][1];
                        uint8_t p8[64], r8[64];
                        orc_mc_luma(RY[0], ys, bx + qx, by + qy, 8, 8, qmv[q][0], qmv[q][1], p8, 8);
                        residual(cy, w, bx + qx, by + qy, p8, 8, qp, lambda, r8, &cod8[q], &d8, &b8);
                        for (int i = 0; i < 64; ++i) rec8[(qy + (i >> 3)) * 16 + qx + (i & 7)] = r8[i];
                    }
                    J8 = d8 + lambda * b8;
                }
            }

            /* commit: strict < in order SKIP, INTER16, INTRA, INTER8 (fdv.h:6036-6040) */
            int mode = 0; double Jm = Js;
            if (gv) mode = gv->mode;
            else {
                if (Ji < Jm) { Jm = Ji; mode = 1; }
                if (Jn < Jm) { Jm = Jn; mode = 2; }
                if (J8 < Jm) { Jm = J8; mode = 3; }
            }
            STRUCT(ctx, mode);
            mmap[my * mw + mx] = (uint8_t)mode;
            if (mode == 0) {
                if (ncand > 1) STRUCT(ORC_S_MERGE, bmi);
                for (int i = 0; i < 16; ++i) memcpy(ry + (by + i) * w + bx, predS + 16 * i, 16);
                int sr = cand[bmi].ref, smx = cand[bmi].x, smy = cand[bmi].y;
                chroma_inter(cu, cw, ru, RU[sr], cs, cbx, cby, smx / 2, smy / 2, qp, 1, &S);
                chroma_inter(cv, cw, rv, RV[sr], cs, cbx, cby, smx / 2, smy / 2, qp, 1, &S);
                mvm[my * mw + mx] = (mvrec_t){(int16_t)smx, (int16_t)smy, (uint8_t)sr, 1};
                pmx = smx; pmy = smy;
                if (mi) *mi = (orc_mbinfo){0, (uint8_t)sr, (uint8_t)bmi, 0, {{(int16_t)smx, (int16_t)smy}}};
            } else if (mode == 1) {
                STRUCT(ORC_S_REF, ibest_ref);
                uint8_t t[10]; size_t tn = leb_write(t, 0, zmap(imx - pmx));
                for (size_t k = 0; k < tn; ++k) STRUCT(ORC_S_MVX, t[k]);
                tn = leb_write(t, 0, zmap(imy - pmy));
                for (size_t k = 0; k < tn; ++k) STRUCT(ORC_S_MVY, t[k]);
                emit_coded(&S, &codI, 0);
                for (int i = 0; i < 16; ++i) memcpy(ry + (by + i) * w + bx, recI + 16 * i, 16);
                pmx = imx; pmy = imy;
                mvm[my * mw + mx] = (mvrec_t){(int16_t)imx, (int16_t)imy, (uint8_t)ibest_ref, 1};
                chroma_inter(cu, cw, ru, RU[ibest_ref], cs, cbx, cby, imx / 2, imy / 2, qp, 0, &S);
                chroma_inter(cv, cw, rv, RV[ibest_ref], cs, cbx, cby, imx / 2, imy / 2, qp, 0, &S);
                if (mi) *mi = (orc_mbinfo){1, (uint8_t)ibest_ref, 0, 0, {{(int16_t)imx, (int16_t)imy}}};
            } else if (mode == 2) {
                STRUCT(ORC_S_SUB, nsub);
                emit_coded(&S, &codN, 0);
                for (int i = 0; i < 16; ++i) memcpy(ry + (by + i) * w + bx, recN + 16 * i, 16);
                const uint8_t *src[2] = {cu, cv};
                uint8_t *dst[2] = {ru, rv};
                for (int pl = 0; pl < 2; ++pl) {
                    uint8_t ct[8], cl[8], ctl, cp[64], cr[64]; int cht, chl;
                    neighbours(dst[pl], cw, cbx, cby, 8, 0, ct, cl, &ctl, &cht, &chl);
                    orc_intra_nxn(nsub, ct, cl, ctl, 8, cht, chl, cp);
                    coded_t cc; double dd = 0.0; int bb = 0;
                    residual(src[pl], cw, cbx, cby, cp, 8, qp, lambda, cr, &cc, &dd, &bb);
                    emit_coded(&S, &cc, 1);
                    for (int i = 0; i < 64; ++i) dst[pl][(cby + (i >> 3)) * cw + cbx + (i & 7)] = cr[i];
                }
                mvm[my * mw + mx] = (mvrec_t){0, 0, 0, 0};
                if (mi) *mi = (orc_mbinfo){2, 0, 0, (uint8_t)nsub, {{0, 0}}};
            } else {
                for (int k = 0; k < s8n; ++k) STRUCT(s8ctx[k], s8[k]);
                for (int q = 0; q < 4; ++q) emit_coded(&S, &cod8[q], 0);
                for (int i = 0; i < 16; ++i) memcpy(ry + (by + i) * w + bx, rec8 + 16 * i, 16);
                pmx = qmv[3][0]; pmy = qmv[3][1];
                mvm[my * mw + mx] = (mvrec_t){(int16_t)qmv[3][0], (int16_t)qmv[3][1], 0, 1};
                const uint8_t *src[2] = {cu, cv};
                uint8_t *dst[2] = {ru, rv};
                const uint8_t *rf[2] = {RU[0], RV[0]};
                for (int pl = 0; pl < 2; ++pl)
                    for (int q = 0; q < 4; ++q) {
                        int qcx = (q & 1) * 4, qcy = (q >> 1) * 4;
                        uint8_t cp[16], cr[16];
                        orc_mc_chroma(rf[pl], cs, cbx + qcx, cby + qcy, 4, 4, qmv[q][0] / 2, qmv[q][1] / 2, cp, 4);
                        coded_t cc; double dd = 0.0; int bb = 0;
                        residual(src[pl], cw, cbx + qcx, cby + qcy, cp, 4, qp, lambda, cr, &cc, &dd, &bb);
                        emit_coded(&S, &cc, 1);
                        for (int i = 0; i < 16; ++i) dst[pl][(cby + qcy + (i >> 2)) * cw + cbx + qcx + (i & 3)] = cr[i];
                    }
                if (mi) {
                    *mi = (orc_mbinfo){3, 0, 0, 0, {{0, 0}}};
                    for (int q = 0; q < 4; ++q) { mi->mv[q][0] = (int16_t)qmv[q][0]; mi->mv[q][1] = (int16_t)qmv[q][1]; }
                }
            }
        }
    }
#undef STRUCT
    if (!orc_no_deblock) {
        orc_deblock_plane(ry, w, h, w, qp);                /* fdv.h:6168-6170 */
        orc_deblock_plane(ru, cw, h / 2, cw, qp);
        orc_deblock_plane(rv, cw, h / 2, cw, qp);
    }

    /* Structure symbols partitioned by model, raster order kept (fdv.h:6180-6193). */
    for (int c = 0; c < ORC_NSTRUCT; ++c) {
        for (size_t i = 0; i < sp; ++i)
            if (st_ctx[i] == c) put_byte(&S, c, st_sym[i]);
    }
    for (int k = 0; k < ORC_NS; ++k) sn[k] = S.n[k];
    free(mmap); free(mvm); free(st_sym); free(st_ctx);
    return S.over ? -1 : 0;
}

int orc_pframe_analyse(const uint8_t *cy, const uint8_t *cu, const uint8_t *cv,
                       int w, int h,
                       const uint8_t *const RY[2], const uint8_t *const RU[2],
                       const uint8_t *const RV[2], int ys, int cs,
                       int navail, int qp, double lambda, int force_skip, int me_mode,
                       const orc_mbinfo *given,
                       uint8_t *ry, uint8_t *ru, uint8_t *rv, orc_mbinfo *info,
                       uint8_t *const bufs[ORC_NS], size_t scap, size_t sn[ORC_NS]) {
    return analyse_core(cy, cu, cv, w, h, RY, RU, RV, ys, cs, navail, qp, lambda, force_skip, me_mode, NULL,
                        given, ry, ru, rv, info, bufs, scap, sn);
}

int orc_pframe_analyse_table(const uint8_t *cy, const uint8_t *cu, const uint8_t *cv,
                             int w, int h,
                             const uint8_t *const RY[2], const uint8_t *const RU[2],
                             const uint8_t *const RV[2], int ys, int cs,
                             int navail, int qp, double lambda, const orc_metab *mt,
                             uint8_t *ry, uint8_t *ru, uint8_t *rv, orc_mbinfo *info,
                             uint8_t *const bufs[ORC_NS], size_t scap, size_t sn[ORC_NS]) {
    return analyse_core(cy, cu, cv, w, h, RY, RU, RV, ys, cs, navail, qp, lambda, 0, ORC_ME_EXHAUSTIVE, mt,
                        NULL, ry, ru, rv, info, bufs, scap, sn);
}

/* ------------------------------------------------------------- entropy -- */

/* Normalise a histogram to 4096 with every used symbol kept >= 1 and the
 * largest bin (first on ties) absorbing the remainder (fdv.h:1804-1840). */
static void freq_norm(const uint32_t cnt[256], uint16_t f[256]) {
    uint64_t tot = 0;
    for (int s = 0; s < 256; ++s) tot += cnt[s];
    memset(f, 0, 256 * sizeof *f);
    if (!tot) { f[0] = 4096; return; }
    uint32_t sum = 0; int big = 0;
    for (int s = 0; s < 256; ++s) {
        if (!cnt[s]) continue;
        uint32_t v = (uint32_t)((uint64_t)cnt[s] * 4096u / tot);
        if (!v) v = 1;
        f[s] = (uint16_t)v;
        sum += v;
        if (v > f[big]) big = s;
    }
    f[big] = (uint16_t)((int)f[big] + (4096 - (int)sum));
}

/* Byte rANS, 12-bit probabilities, state in [2^23, 2^31), symbols encoded
 * last to first, 4-byte little-endian final state first in the output
 * (fdv.h:1760-1801, 1936-1948). */
static size_t rans_enc(const uint8_t *sym, size_t n, const uint16_t f[256], uint8_t *out, size_t cap) {
    uint16_t cum[256]; uint32_t c = 0;
    for (int s = 0; s < 256; ++s) { cum[s] = (uint16_t)c; c += f[s]; }
    uint8_t *p = out + cap;
    uint32_t x = 1u << 23;
    for (size_t i = n; i-- > 0;) {
        uint32_t fs = f[sym[i]];
        uint64_t xmax = (uint64_t)((1u << 23) >> 12 << 8) * fs;
        while (x >= xmax) { *--p = (uint8_t)x; x >>= 8; }
        x = ((x / fs) << 12) + (x % fs) + cum[sym[i]];
    }
    p -= 4;
    p[0] = (uint8_t)x; p[1] = (uint8_t)(x >> 8); p[2] = (uint8_t)(x >> 16); p[3] = (uint8_t)(x >> 24);
    size_t len = (size_t)(out + cap - p);
    memmove(out, p, len);
    return len;
}

/* Table wire form: u16 count, then (symbol, LEB freq) pairs (fdv.h:1862-1886). */
static size_t freqs_size(const uint16_t f[256]) {
    size_t n = 2;
    for (int s = 0; s < 256; ++s) if (f[s]) n += 1 + (size_t)leb_len(f[s]);
    return n;
}
static size_t freqs_write(uint8_t *o, size_t p, const uint16_t f[256]) {
    int n = 0;
    for (int s = 0; s < 256; ++s) n += f[s] != 0;
    o[p++] = (uint8_t)n; o[p++] = (uint8_t)(n >> 8);
    for (int s = 0; s < 256; ++s) if (f[s]) { o[p++] = (uint8_t)s; p = leb_write(o, p, f[s]); }
    return p;
}
static int covers(const uint16_t f[256], const uint32_t c[256]) {     /* fdv.h:5641 */
    for (int s = 0; s < 256; ++s) if (c[s] && !f[s]) return 0;
    return 1;
}
/* The cached form of a table: unused symbols get one slot, paid by the
 * largest bin, unless it cannot spare them (tab_store, fdv.h:5667-5677). */
static void store_tab(uint16_t d[256], const uint16_t s[256]) {
    int add = 0, big = 0;
    for (int i = 0; i < 256; ++i) { add += !s[i]; if (s[i] > s[big]) big = i; }
    if (s[big] <= add) { memcpy(d, s, 256 * sizeof *d); return; }
    for (int i = 0; i < 256; ++i) d[i] = s[i] ? s[i] : 1;
    d[big] = (uint16_t)(d[big] - add);
}
static void hist_fold(uint32_t h[256], const uint32_t c[256]) {       /* fdv.h:5661 */
    uint64_t s = 0;
    for (int i = 0; i < 256; ++i) { h[i] += c[i]; s += h[i]; }
    if (s > (1u << 21)) for (int i = 0; i < 256; ++i) h[i] >>= 1;
}
/* Streams coded adaptively: shortest first (stable), within 65536 symbols
 * (ad_select, fdv.h:482-499). */
static unsigned ada_pick(const size_t sn[ORC_NS], size_t *tot) {
    int ord[ORC_NS];
    for (int i = 0; i < ORC_NS; ++i) ord[i] = i;
    for (int i = 1; i < ORC_NS; ++i)
        for (int j = i; j > 0 && sn[ord[j]] < sn[ord[j - 1]]; --j) { int t = ord[j]; ord[j] = ord[j - 1]; ord[j - 1] = t; }
    unsigned m = 0; size_t t = 0;
    for (int i = 0; i < ORC_NS; ++i) {
        int k = ord[i];
        if (!sn[k] || t + sn[k] > 65536) continue;
        m |= 1u << k; t += sn[k];
    }
    *tot = t;
    return m;
}

/* Adaptive model: weights 1 + history (normalised to 8192 when larger), +128
 * per coded symbol, halved (rounding up) when the total passes 65536
 * (fdv_amodel_init / am_add, fdv.h:2024-2063).  Cumulative sums by plain
 * summation -- the Fenwick tree of the reference is only an accelerator. */
typedef struct { uint32_t f[256]; uint32_t tot; } amod_t;
static void amod_init(amod_t *m, const uint32_t h[256]) {
    uint64_t s = 0;
    for (int i = 0; i < 256; ++i) s += h[i];
    m->tot = 0;
    for (int i = 0; i < 256; ++i) {
        m->f[i] = 1 + (s <= 8192 ? h[i] : (uint32_t)(((uint64_t)h[i] * 8192 + s / 2) / s));
        m->tot += m->f[i];
    }
}
static void amod_bump(amod_t *m, unsigned s) {
    m->f[s] += 128; m->tot += 128;
    if (m->tot > 65536) {
        m->tot = 0;
        for (int i = 0; i < 256; ++i) { m->f[i] = (m->f[i] + 1) >> 1; m->tot += m->f[i]; }
    }
}
/* Carry-propagating range coder, 32-bit range, byte output with a cache byte
 * and a pending-0xFF run (fdv.h:2082-2118). */
typedef struct { uint8_t *o; size_t cap, n; uint64_t low; uint32_t range; uint8_t cache; uint64_t pend; int fail; } rc_t;
static void rc_shift(rc_t *e) {
    if ((uint32_t)e->low < 0xFF000000u || (e->low >> 32)) {
        uint8_t carry = (uint8_t)(e->low >> 32), c = e->cache;
        do {
            if (e->n < e->cap) e->o[e->n++] = (uint8_t)(c + carry); else e->fail = 1;
            c = 0xFF;
        } while (--e->pend);
        e->cache = (uint8_t)((uint32_t)e->low >> 24);
    }
    e->pend++;
    e->low = (uint32_t)e->low << 8;
}
static void rc_put(rc_t *e, amod_t *m, unsigned s) {
    uint32_t cum = 0;
    for (unsigned i = 0; i < s; ++i) cum += m->f[i];
    uint32_t r = e->range / m->tot;
    e->low += (uint64_t)r * cum;
    e->range = r * m->f[s];
    while (e->range < (1u << 24)) { rc_shift(e); e->range <<= 8; }
    amod_bump(m, s);
}
static size_t rc_done(rc_t *e) {
    for (int i = 0; i < 5; ++i) rc_shift(e);
    if (e->fail || e->n < 1) return 0;
    memmove(e->o, e->o + 1, e->n - 1);      /* the always-zero first byte */
    return e->n - 1;
}

/* Four candidate codings, smallest wins (ties to the lower mode number):
 * 0 one fresh shared table, 1 a table per stream (fresh or reused, per-stream
 * mask), 2 the cached shared table, 3 as 1 with the short streams coded
 * adaptively from history (fdv.h:6172-6421). */
size_t orc_pframe_pack(int qp, const uint8_t *const sv[ORC_NS], const size_t sn[ORC_NS],
                       orc_tabcache *tc, uint8_t *out, size_t cap) {
    size_t total = 0;
    for (int k = 0; k < ORC_NS; ++k) total += sn[k];
    uint8_t *all = malloc(total ? total : 1);
    uint32_t (*ck)[256] = calloc(ORC_NS, sizeof *ck);
    uint32_t c1[256] = {0};
    uint16_t f1[256], (*fk)[256] = calloc(ORC_NS, sizeof *fk);
    size_t bufcap = total * 2 + 1024;
    uint8_t *p1 = malloc(bufcap), *p1r = NULL, *pada = NULL;
    uint8_t *pay[ORC_NS] = {0}, *payr[ORC_NS] = {0};
    size_t len[ORC_NS] = {0}, lenr[ORC_NS] = {0};
    size_t result = 0;
    if (!all || !ck || !fk || !p1) goto done;
    {
        size_t o = 0;
        for (int k = 0; k < ORC_NS; ++k) {
            memcpy(all + o, sv[k], sn[k]); o += sn[k];
            for (size_t i = 0; i < sn[k]; ++i) ck[k][sv[k][i]]++;
        }
    }
    for (size_t i = 0; i < total; ++i) c1[all[i]]++;
    freq_norm(c1, f1);
    size_t l1 = rans_enc(all, total, f1, p1, bufcap);
    for (int k = 0; k < ORC_NS; ++k) {
        freq_norm(ck[k], fk[k]);
        pay[k] = malloc(sn[k] * 2 + 1024);
        if (!pay[k]) goto done;
        len[k] = rans_enc(sv[k], sn[k], fk[k], pay[k], sn[k] * 2 + 1024);
    }
    size_t hdr = 0;
    for (int k = 0; k < ORC_NS; ++k) hdr += (size_t)leb_len((uint32_t)sn[k]);
    size_t size_one = hdr + freqs_size(f1) + 4 + l1;

    size_t l1r = 0, size_one_r = (size_t)-1, size_two_r = (size_t)-1;
    if (tc && tc->have) {
        if (covers(tc->f1, c1) && (p1r = malloc(bufcap)) != NULL) {
            l1r = rans_enc(all, total, tc->f1, p1r, bufcap);
            if (l1r) size_one_r = hdr + 4 + l1r;
        }
        int all_cover = 1;
        for (int k = 0; k < ORC_NS; ++k) all_cover &= covers(tc->fs[k], ck[k]);
        if (all_cover) {
            int ok = 1;
            for (int k = 0; k < ORC_NS && ok; ++k) {
                payr[k] = malloc(sn[k] * 2 + 1024);
                if (!payr[k]) { ok = 0; break; }
                lenr[k] = rans_enc(sv[k], sn[k], tc->fs[k], payr[k], sn[k] * 2 + 1024);
            }
            if (ok) { size_two_r = hdr; for (int k = 0; k < ORC_NS; ++k) size_two_r += 4 + lenr[k]; }
        }
    }
    int have_reuse = tc && tc->have && size_two_r != (size_t)-1;
    size_t size_mix = hdr + 2;
    unsigned mask = 0;
    for (int k = 0; k < ORC_NS; ++k) {
        if (!sn[k]) continue;
        size_t fresh = freqs_size(fk[k]) + (size_t)leb_len((uint32_t)len[k]) + len[k];
        size_t reused = have_reuse ? (size_t)leb_len((uint32_t)lenr[k]) + lenr[k] : (size_t)-1;
        if (reused < fresh) { size_mix += reused; mask |= 1u << k; } else size_mix += fresh;
    }
    size_t size_ada = (size_t)-1, lada = 0;
    unsigned amask = 0;
    if (tc && tc->hn) {
        size_t asyms = 0;
        amask = ada_pick(sn, &asyms);
        if (amask && (pada = malloc(asyms * 3 + 64)) != NULL) {
            rc_t e = {pada, asyms * 3 + 64, 0, 0, 0xFFFFFFFFu, 0, 1, 0};
            for (int k = 0; k < ORC_NS; ++k) {
                if (!((amask >> k) & 1)) continue;
                amod_t m; amod_init(&m, tc->hist[k]);
                for (size_t i = 0; i < sn[k]; ++i) rc_put(&e, &m, sv[k][i]);
            }
            lada = rc_done(&e);
            if (lada) {
                size_ada = hdr + 2 + (size_t)leb_len((uint32_t)lada) + lada;
                for (int k = 0; k < ORC_NS; ++k) {
                    if (!sn[k] || ((amask >> k) & 1)) continue;
                    size_t fresh = freqs_size(fk[k]) + (size_t)leb_len((uint32_t)len[k]) + len[k];
                    size_t reused = have_reuse ? (size_t)leb_len((uint32_t)lenr[k]) + lenr[k] : (size_t)-1;
                    size_ada += reused < fresh ? reused : fresh;
                }
            }
        }
    }
    int emode = 0; size_t best = size_one;
    if (size_mix < best) { best = size_mix; emode = 1; }
    if (size_one_r < best) { best = size_one_r; emode = 2; }
    if (size_ada < best) { best = size_ada; emode = 3; }
    if (2 + best > cap) goto done;
    {
        size_t o = 0;
        out[o++] = (uint8_t)qp;
        out[o++] = (uint8_t)emode;
        for (int k = 0; k < ORC_NS; ++k) o = leb_write(out, o, (uint32_t)sn[k]);
        if (!(emode & 1)) {
            if (!(emode & 2)) o = freqs_write(out, o, f1);
            size_t n = (emode & 2) ? l1r : l1;
            const uint8_t *q = (emode & 2) ? p1r : p1;
            for (int i = 0; i < 4; ++i) out[o++] = (uint8_t)(n >> (8 * i));
            memcpy(out + o, q, n); o += n;
        } else {
            unsigned am = emode == 3 ? amask : 0u, sm = mask & ~am;
            out[o++] = (uint8_t)sm; out[o++] = (uint8_t)(sm >> 8);
            for (int k = 0; k < ORC_NS; ++k) {
                if (!sn[k] || ((am >> k) & 1)) continue;
                int reused = (mask >> k) & 1;
                if (!reused) o = freqs_write(out, o, fk[k]);
                size_t n = reused ? lenr[k] : len[k];
                o = leb_write(out, o, (uint32_t)n);
                memcpy(out + o, reused ? payr[k] : pay[k], n); o += n;
            }
            if (emode == 3) { o = leb_write(out, o, (uint32_t)lada); memcpy(out + o, pada, lada); o += lada; }
        }
        if (tc) {
            if (emode == 0) { store_tab(tc->f1, f1); tc->have = 1; }
            else if (emode == 1 || emode == 3) {
                unsigned am = emode == 3 ? amask : 0u;
                for (int k = 0; k < ORC_NS; ++k)
                    if (sn[k] && !((am >> k) & 1) && !((mask >> k) & 1)) store_tab(tc->fs[k], fk[k]);
                tc->have = 1;
            }
            for (int k = 0; k < ORC_NS; ++k) hist_fold(tc->hist[k], ck[k]);
            if (tc->hn < (1 << 20)) ++tc->hn;
        }
        result = o;
    }
done:
    free(all); free(ck); free(fk); free(p1); free(p1r); free(pada);
    for (int k = 0; k < ORC_NS; ++k) { free(pay[k]); free(payr[k]); }
    return result;
}

/* ------------------------------------------------------------ wrapper -- */

size_t orc_pframe_encode(const uint8_t *cy, const uint8_t *cu, const uint8_t *cv,
                         int w, int h,
                         const uint8_t *r0y, const uint8_t *r0u, const uint8_t *r0v,
                         const uint8_t *r1y, const uint8_t *r1u, const uint8_t *r1v,
                         int navail, int qp, uint8_t *out, size_t cap,
                         uint8_t *ry, uint8_t *ru, uint8_t *rv, int force_skip,
                         orc_tabcache *tc, double lscale, int me_mode,
                         orc_mbinfo *mbinfo) {
    int ys = orc_plane_stride(w), cs = orc_plane_stride(w / 2);
    size_t ysz = (size_t)ys * (h + 2 * ORC_BORDER), csz = (size_t)cs * (h / 2 + 2 * ORC_BORDER);
    uint8_t *mem = malloc(2 * (ysz + 2 * csz));
    size_t nmb = (size_t)(w / 16) * (h / 16);
    size_t scap = nmb * 1600 + 1024;
    uint8_t *sb = malloc(scap * ORC_NS);
    if (!mem || !sb) { free(mem); free(sb); return 0; }
    const uint8_t *srcs[2][3] = {{r0y, r0u, r0v}, {r1y ? r1y : r0y, r1u ? r1u : r0u, r1v ? r1v : r0v}};
    const uint8_t *RY[2], *RU[2], *RV[2];
    uint8_t *m = mem;
    for (int r = 0; r < 2; ++r) {
        uint8_t *py = m + (size_t)ORC_BORDER * ys + ORC_BORDER; m += ysz;
        uint8_t *pu = m + (size_t)ORC_BORDER * cs + ORC_BORDER; m += csz;
        uint8_t *pv = m + (size_t)ORC_BORDER * cs + ORC_BORDER; m += csz;
        orc_plane_from_tight(py, ys, srcs[r][0], w, h);
        orc_plane_from_tight(pu, cs, srcs[r][1], w / 2, h / 2);
        orc_plane_from_tight(pv, cs, srcs[r][2], w / 2, h / 2);
        RY[r] = py; RU[r] = pu; RV[r] = pv;
    }
    if (navail < 2) { RY[1] = RY[0]; RU[1] = RU[0]; RV[1] = RV[0]; }
    uint8_t *bufs[ORC_NS];
    size_t sn[ORC_NS];
    for (int k = 0; k < ORC_NS; ++k) bufs[k] = sb + (size_t)k * scap;
    size_t n = 0;
    double lambda = orc_lambda(qp, lscale);
    if (orc_pframe_analyse(cy, cu, cv, w, h, RY, RU, RV, ys, cs, navail, qp, lambda, force_skip,
                           me_mode, NULL, ry, ru, rv, mbinfo, bufs, scap, sn) == 0)
        n = orc_pframe_pack(qp, (const uint8_t *const *)bufs, sn, tc, out, cap);
    free(mem); free(sb);
    return n;
}


/* ========================================================= motion tables ==
 * The per-MB motion tables of the library's exhaustive, parallel and learned
 * modes (not in the reference): the vectors and SADs the MB loop would get
 * from its searches, computed for every MB at once.  The exhaustive search
 * ignores the predictor, so its table is a pure function of the frame and the
 * references; orc_pframe_analyse_table runs the MB loop on a table. */

void orc_metab_exhaustive(const uint8_t *cy, int w, int h, const uint8_t *const RY[2], int ys,
                          int navail, orc_metab *mt) {
    const int mw = w / 16, mh = h / 16;
    for (int my = 0; my < mh; ++my)
        for (int mx = 0; mx < mw; ++mx) {
            orc_metab *m = &mt[my * mw + mx];
            memset(m, 0, sizeof *m);
            const int bx = mx * 16, by = my * 16;
            for (int r = 0; r < navail; ++r) {
                int a, b;
                m->sad16[r] = orc_me_search(cy, w, RY[r], ys, bx, by, 16, 16, 16, 0, 0, ORC_ME_EXHAUSTIVE, &a, &b);
                m->mv16[r][0] = (int16_t)a; m->mv16[r][1] = (int16_t)b;
            }
            for (int q = 0; q < 4; ++q) {
                int a, b;
                m->sad8[q] = orc_me_search(cy, w, RY[0], ys, bx + (q & 1) * 8, by + (q >> 1) * 8, 8, 8, 16,
                                           0, 0, ORC_ME_EXHAUSTIVE, &a, &b);
                m->mv8[q][0] = (int16_t)a; m->mv8[q][1] = (int16_t)b;
            }
            if (navail < 2) {                   /* the unused far entry repeats the near one */
                m->mv16[1][0] = m->mv16[0][0]; m->mv16[1][1] = m->mv16[0][1]; m->sad16[1] = m->sad16[0];
            }
        }
}

/* Whole-pel search over a point set with the exhaustive search's tie order
 * (SAD, then |dx|+|dy|, then raster index over the +-range grid), then the
 * reference's sub-pel cross. */
static int search_points(const uint8_t *cur, int cs, const uint8_t *ref, int rs, int bx, int by, int bs,
                         const int (*pts)[2], int npts, int *mvx, int *mvy) {
    long long bestk = -1;
    int bxv = 0, byv = 0, best = 0;
    for (int k = 0; k < npts; ++k) {
        const int dx = pts[k][0], dy = pts[k][1];
        const int s = sad_at(cur, cs, ref, rs, bx, by, bs, bs, dx * 4, dy * 4);
        const long long key = ((long long)s << 32) | ((long long)(iabs(dx) + iabs(dy)) << 16) | ((dy + 16) * 33 + dx + 16);
        if (bestk < 0 || key < bestk) { bestk = key; best = s; bxv = dx * 4; byv = dy * 4; }
    }
    subpel_refine(cur, cs, ref, rs, bx, by, bs, bs, &bxv, &byv, &best);
    *mvx = bxv; *mvy = byv;
    return best;
}

void orc_metab_candidates(const uint8_t *cy, int w, int h, const uint8_t *const RY[2], int ys, int navail,
                          const orc_nnout *nn, float thr, float sthr, orc_metab *mt) {
    const int mw = w / 16, mh = h / 16;
    orc_metab_exhaustive(cy, w, h, RY, ys, navail, mt);     /* the fallback for every MB */
    for (int i = 0; i < mw * mh; ++i) {
        const orc_nnout *P = &nn[i];
        orc_metab *m = &mt[i];
        m->flags = P->pskip >= sthr ? 1 : 0;
        if (!(P->conf >= thr)) continue;
        const int bx = (i % mw) * 16, by = (i / mw) * 16;
        for (int r = 0; r < navail; ++r) {
            int pts[181][2], n = 0;
            for (int c = 0; c < (r ? 12 : 20); ++c)     /* the quadrant maps' 8 on the near reference only */
                for (int o = 0; o < 9; ++o) {
                    const int vx = c < 8 ? P->cx[c] : c < 12 ? P->kx[c - 8] : P->qx[(c - 12) >> 1][(c - 12) & 1];
                    const int vy = c < 8 ? P->cy[c] : c < 12 ? P->ky[c - 8] : P->qy[(c - 12) >> 1][(c - 12) & 1];
                    const int dx = (r + 1) * vx + o % 3 - 1, dy = (r + 1) * vy + o / 3 - 1;
                    if (dx < -16 || dx > 16 || dy < -16 || dy > 16) continue;
                    pts[n][0] = dx; pts[n][1] = dy; ++n;
                }
            pts[n][0] = 0; pts[n][1] = 0; ++n;
            int a, b;
            m->sad16[r] = search_points(cy, w, RY[r], ys, bx, by, 16, (const int (*)[2])pts, n, &a, &b);
            m->mv16[r][0] = (int16_t)a; m->mv16[r][1] = (int16_t)b;
            if (r == 0)
                for (int q = 0; q < 4; ++q) {
                    m->sad8[q] = search_points(cy, w, RY[0], ys, bx + (q & 1) * 8, by + (q >> 1) * 8, 8,
                                               (const int (*)[2])pts, n, &a, &b);
                    m->mv8[q][0] = (int16_t)a; m->mv8[q][1] = (int16_t)b;
                }
        }
        if (navail < 2) { m->mv16[1][0] = m->mv16[0][0]; m->mv16[1][1] = m->mv16[0][1]; m->sad16[1] = m->sad16[0]; }
    }
}

/* ------------------------------------------------------- the key frame -- */

/* The intra-only plane codec (fdv_image_encode, fdv.h:4845-5038), restated as
 * two separate passes: a decision pass that walks the rate-distortion quadtree
 * and records the winner of every node in position-indexed planes (split
 * flags per depth, leaf modes, region flags, levels, reconstruction), and a
 * serialising pass that walks the decided tree in coding order and emits the
 * twelve streams.  The reference emits while it decides, keeping the
 * symbols of both trials of a node until one wins; the stream contents are
 * the same because a stream holds exactly the winners' symbols in pre-order. */

#define ICTU 64                                   /* FDV_CTU (fdv.h:3651) */
enum { IS_SPLIT = 0, IS_MODEB = 4, IS_MODE0, IS_MODE1, IS_CNT, IS_LVL, IS_CNT8, IS_LVL8, IS_FLAG, IS_NS };

/* Z-order rank of 4x4 cell (lx, ly) inside a tree unit: vertical bit above the
 * horizontal one at each level (zidx4, fdv.h:3659). */
static int izrank(int lx, int ly) {
    int z = 0;
    for (int b = ICTU / 8; b; b >>= 1) z = (z << 2) | ((ly & b) ? 2 : 0) | ((lx & b) ? 1 : 0);
    return z;
}
/* Sample (x, y) is reconstructed while the 4x4 block at (bx, by) is coded:
 * an earlier unit row, an earlier unit of this row, or an earlier cell of
 * this unit in z-order (cell_done, fdv.h:3677). */
static int idone(int x, int y, int bx, int by) {
    int ux = bx & ~(ICTU - 1), uy = by & ~(ICTU - 1);
    if (y < uy) return 1;
    if (x < ux) return 1;
    if (x >= ux + ICTU) return 0;
    return izrank((x - ux) >> 2, (y - uy) >> 2) < izrank((bx - ux) >> 2, (by - uy) >> 2);
}

/* The nine 4x4 predictors of H.264 (8.3.1.2), with the edge laid out as one
 * array: E[3 - i] = left sample i, E[4] = the corner, E[5 + i] = top sample i
 * (i = 0..7, top-right included).  DC follows the availability flags
 * (fdv_intra_predict_4x4, fdv.h:2698). */
static int f3(int a, int b, int c) { return (a + 2 * b + c + 2) >> 2; }
static int f2(int a, int b) { return (a + b + 1) >> 1; }
void orc_intra4(int mode, const uint8_t top[8], const uint8_t left[4], uint8_t tl, int ht, int hl,
                uint8_t pred[16]) {
    int E[13];
    for (int i = 0; i < 4; ++i) E[3 - i] = left[i];
    E[4] = tl;
    for (int i = 0; i < 8; ++i) E[5 + i] = top[i];
#define TP(i) E[5 + (i)]
#define LF(i) E[3 - (i)]
    for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x) {
            int v;
            switch (mode) {
            case 1: v = TP(x); break;                                    /* vertical */
            case 2: v = LF(y); break;                                    /* horizontal */
            case 3: v = (x == 3 && y == 3) ? (TP(6) + 3 * TP(7) + 2) >> 2   /* diagonal down-left */
                                           : f3(TP(x + y), TP(x + y + 1), TP(x + y + 2)); break;
            case 4: v = f3(E[3 + x - y], E[4 + x - y], E[5 + x - y]); break;   /* diagonal down-right */
            case 5: {                                                    /* vertical-right */
                int z = 2 * x - y, k = x - (y >> 1);
                if (z >= 0) v = (z & 1) ? f3(E[3 + k], E[4 + k], E[5 + k]) : f2(E[4 + k], E[5 + k]);
                else if (z == -1) v = f3(E[3], E[4], E[5]);
                else v = f3(E[4 - y], E[5 - y], E[6 - y]);
                break;
            }
            case 6: {                                                    /* horizontal-down */
                int z = 2 * y - x, k = y - (x >> 1);
                if (z >= 0) v = (z & 1) ? f3(E[5 - k], E[4 - k], E[3 - k]) : f2(E[4 - k], E[3 - k]);
                else if (z == -1) v = f3(E[5], E[4], E[3]);
                else v = f3(E[4 + x], E[3 + x], E[2 + x]);
                break;
            }
            case 7: {                                                    /* vertical-left */
                int k = x + (y >> 1);
                v = (y & 1) ? f3(TP(k), TP(k + 1), TP(k + 2)) : f2(TP(k), TP(k + 1));
                break;
            }
            case 8: {                                                    /* horizontal-up */
                int z = x + 2 * y, k = y + (x >> 1);
                if (z < 5) v = (z & 1) ? f3(LF(k), LF(k + 1), LF(k + 2)) : f2(LF(k), LF(k + 1));
                else if (z == 5) v = (LF(2) + 3 * LF(3) + 2) >> 2;
                else v = LF(3);
                break;
            }
            default: {                                                   /* DC */
                int s = 0;
                if (ht && hl) { for (int i = 0; i < 4; ++i) s += top[i] + left[i]; v = (s + 4) >> 3; }
                else if (ht) { for (int i = 0; i < 4; ++i) s += top[i]; v = (s + 2) >> 2; }
                else if (hl) { for (int i = 0; i < 4; ++i) s += left[i]; v = (s + 2) >> 2; }
                else v = 128;
            }
            }
            pred[y * 4 + x] = (uint8_t)v;
        }
#undef TP
#undef LF
}

/* Which 4x4 modes their neighbours allow (mode_ok, fdv.h:3709). */
static int m4_ok(int m, int ht, int htr, int hl, int htl) {
    switch (m) {
    case 0: return 1;
    case 1: return ht;
    case 2: case 8: return hl;
    case 3: case 7: return ht && htr;
    default: return ht && hl && htl;
    }
}

typedef struct {
    const uint8_t *src; int ss, w, h, qp; double lambda;
    orc_itree *t;
} iplane_t;

static int inodes_w(int w, int d) { int n = ICTU >> d; return (w + n - 1) / n; }

/* A 4x4 leaf: the SATD screen over the allowed modes, the two best (all of
 * them below QP 12) coded without RDOQ and compared in D + lambda*R, pruned
 * once a candidate's SATD passes 5/4 of the best; the winner recoded with
 * RDOQ (intra_leaf4, fdv.h:3907-4108).  Returns the winner's cost. */
static double ileaf4(iplane_t *P, int bx, int by) {
    orc_itree *t = P->t;
    const int w = P->w, qp = P->qp;
    const double lambda = P->lambda;
    uint8_t *rec = t->rec;
    const int ht = by > 0, hl = bx > 0, htl = ht && hl;
    const int htr = ht && bx + 4 < w && idone(bx + 4, by - 1, bx, by);
    uint8_t top[8] = {0}, left[4] = {0}, tl = 128;
    if (ht) for (int j = 0; j < 8; ++j) top[j] = rec[(size_t)(by - 1) * w + bx + (j < 4 || htr ? j : 3)];
    if (hl) for (int i = 0; i < 4; ++i) left[i] = rec[(size_t)(by + i) * w + bx - 1];
    if (htl) tl = rec[(size_t)(by - 1) * w + bx - 1];
    const uint8_t *sb = P->src + (size_t)by * P->ss + bx;

    uint8_t pred[9][16];
    int sad[9], md[9], nc = 0;
    for (int m = 0; m < 9; ++m) {
        if (!m4_ok(m, ht, htr, hl, htl)) continue;
        orc_intra4(m, top, left, tl, ht, hl, pred[nc]);
        sad[nc] = orc_satd4(sb, P->ss, pred[nc]);
        md[nc] = m;
        ++nc;
    }
    /* order the first ntry by SATD: selection with exchanges, as the reference
     * does it (the exchange moves the displaced entry, which matters for ties) */
    const int ntry = qp < 12 ? nc : (nc < 2 ? nc : 2);
    int ord[9];
    for (int i = 0; i < nc; ++i) ord[i] = i;
    for (int a = 0; a < ntry; ++a) {
        int bst = a;
        for (int b = a + 1; b < nc; ++b) if (sad[ord[b]] < sad[ord[bst]]) bst = b;
        int tmp = ord[a]; ord[a] = ord[bst]; ord[bst] = tmp;
    }
    int16_t res[16];
    double best = -1.0;
    int bc = 0;
    for (int c = 0; c < ntry; ++c) {
        if (c > 0 && qp >= 12 && (int64_t)sad[ord[c]] * 4 > (int64_t)sad[ord[0]] * 5) break;
        const uint8_t *pp = pred[ord[c]];
        int16_t lv[16], rr[16];
        int32_t co[16], dq[16];
        for (int i = 0; i < 16; ++i) res[i] = (int16_t)(sb[(i >> 2) * P->ss + (i & 3)] - pp[i]);
        orc_fdct4(res, co);
        orc_quant4(co, lv, qp);
        orc_dequant4(lv, dq, qp);
        orc_idct4(dq, rr);
        int32_t e = 0;
        for (int i = 0; i < 16; ++i) { int d = sb[(i >> 2) * P->ss + (i & 3)] - clip8(pp[i] + rr[i]); e += d * d; }
        int last = -1;
        for (int k = 0; k < 16; ++k) if (lv[SCAN4[k]]) last = k;
        int bits = 3 + orc_bits_val((uint32_t)(last + 1));          /* FDV_BITS_MODE (fdv.h:3351) */
        for (int k = 0; k <= last; ++k) bits += orc_bits_val(zmap(lv[SCAN4[k]]));
        double J = (double)e + lambda * bits;
        if (best < 0.0 || J < best) { best = J; bc = c; }
    }
    const uint8_t *pp = pred[ord[bc]];
    int16_t lv[16], rr[16];
    int32_t co[16], dq[16];
    for (int i = 0; i < 16; ++i) res[i] = (int16_t)(sb[(i >> 2) * P->ss + (i & 3)] - pp[i]);
    orc_fdct4(res, co);
    orc_quant4(co, lv, qp);
    orc_rdoq4(lv, res, qp, lambda);
    orc_dequant4(lv, dq, qp);
    orc_idct4(dq, rr);
    int32_t e = 0;
    for (int i = 0; i < 16; ++i) {
        int v = clip8(pp[i] + rr[i]);
        rec[(size_t)(by + (i >> 2)) * w + bx + (i & 3)] = (uint8_t)v;
        t->lev[(size_t)(by + (i >> 2)) * w + bx + (i & 3)] = lv[i];
        int d = sb[(i >> 2) * P->ss + (i & 3)] - v;
        e += d * d;
    }
    int last = -1;
    for (int k = 0; k < 16; ++k) if (lv[SCAN4[k]]) last = k;
    int bits = 3 + orc_bits_val((uint32_t)(last + 1));
    for (int k = 0; k <= last; ++k) bits += orc_bits_val(zmap(lv[SCAN4[k]]));
    t->cmode[(size_t)(by / 4) * (w / 4) + bx / 4] = (uint8_t)md[ord[bc]];
    return (double)e + lambda * bits;
}

/* An n x n leaf (n = 8 .. 64): each allowed DC / V / H / PLANE prediction from
 * the unfiltered neighbours, its residual coded region by region; the first
 * lowest D + lambda*R wins (intra_leafn, fdv.h:4442-4484).  The winner goes to
 * the planes only if the leaf beats the split, so it is returned in buffers. */
static double ileafn(iplane_t *P, int bx, int by, int n, int *mode, uint8_t *brec, int16_t *blev,
                     uint8_t *bflag) {
    const orc_itree *t = P->t;
    const int w = P->w;
    const int ht = by > 0, hl = bx > 0;
    uint8_t top[ICTU], left[ICTU], tl = 128, pred[ICTU * ICTU], trec[ICTU * ICTU], tflag[64];
    static __thread int16_t tlev[ICTU * ICTU];
    for (int k = 0; k < n; ++k) {
        top[k] = ht ? t->rec[(size_t)(by - 1) * w + bx + k] : 0;
        left[k] = hl ? t->rec[(size_t)(by + k) * w + bx - 1] : 0;
    }
    if (ht && hl) tl = t->rec[(size_t)(by - 1) * w + bx - 1];
    double best = -1.0;
    const int nr = n / 8;
    for (int m = 0; m < 4; ++m) {
        if (!nn_ok(m, n, ht, hl)) continue;
        orc_intra_nxn(m, top, left, tl, n, ht, hl, pred);
        double D = 0.0;
        int bits = 3;                                                  /* the mode symbol */
        for (int r = 0; r < nr * nr; ++r) {
            const int rx = (r % nr) * 8, ry = (r / nr) * 8;
            double e; int b;
            region8(P->src + (size_t)(by + ry) * P->ss + bx + rx, P->ss, pred + ry * n + rx, n, P->qp,
                    P->lambda, trec + ry * n + rx, n, &tflag[r], tlev + 64 * r, &e, &b);
            D += e; bits += b;
        }
        double J = D + P->lambda * bits;
        if (best < 0.0 || J < best) {
            best = J; *mode = m;
            memcpy(brec, trec, (size_t)n * n);
            memcpy(blev, tlev, (size_t)n * n * sizeof *blev);
            memcpy(bflag, tflag, (size_t)nr * nr);
        }
    }
    return best;
}

/* A 4x4-mode-context value for a large leaf's cells (nn_to_4x4_mode, fdv.h:4400). */
static int nn_ctx(int m) { return m == 1 ? 1 : m == 2 ? 2 : 0; }

/* One node: implied split off the plane's edge; else the split (children in
 * z-order) against the node as one leaf, the leaf on ties (intra_node,
 * fdv.h:4494-4584).  Returns the winner's cost; the planes hold the winner. */
static double inode(iplane_t *P, int bx, int by, int d) {
    orc_itree *t = P->t;
    const int n = ICTU >> d, hn = n / 2, w = P->w, h = P->h;
    if (n == 4) return ileaf4(P, bx, by);
    const size_t ni = (size_t)(by / n) * inodes_w(w, d) + bx / n;
    if (bx + n > w || by + n > h) {
        t->split[d][ni] = 2;
        double J = 0.0;
        for (int k = 0; k < 4; ++k) {
            int cx = bx + (k & 1) * hn, cy = by + (k >> 1) * hn;
            if (cx < w && cy < h) J += inode(P, cx, cy, d + 1);
        }
        return J;
    }
    double Js = 0.0;
    for (int k = 0; k < 4; ++k) Js += inode(P, bx + (k & 1) * hn, by + (k >> 1) * hn, d + 1);
    uint8_t brec[ICTU * ICTU], bflag[64];
    static __thread int16_t blev[ICTU * ICTU];
    int mode = 0;
    const double Jl = ileafn(P, bx, by, n, &mode, brec, blev, bflag);
    const int split = Js < Jl;
    t->split[d][ni] = (uint8_t)split;
    if (split) return Js;
    t->nmode[d][ni] = (uint8_t)mode;
    const int nr = n / 8, rw = (w + 7) / 8;
    for (int y = 0; y < n; ++y)
        memcpy(t->rec + (size_t)(by + y) * w + bx, brec + y * n, (size_t)n);
    for (int r = 0; r < nr * nr; ++r) {
        const int rx = bx + (r % nr) * 8, ry = by + (r / nr) * 8;
        t->rflag[(size_t)(ry / 8) * rw + rx / 8] = bflag[r];
        for (int i = 0; i < 64; ++i) {
            /* flag 1: four 4x4 blocks of 16 raster levels; else one 8x8 raster */
            const int y = bflag[r] == 1 ? ((i >> 5) << 2) + ((i >> 2) & 3) : i >> 3;
            const int x = bflag[r] == 1 ? (((i >> 4) & 1) << 2) + (i & 3) : i & 7;
            t->lev[(size_t)(ry + y) * w + rx + x] = blev[64 * r + i];
        }
    }
    for (int y = 0; y < n / 4; ++y)
        memset(t->cmode + (size_t)(by / 4 + y) * (w / 4) + bx / 4, nn_ctx(mode), (size_t)n / 4);
    return Jl;
}

void orc_image_analyse(const uint8_t *src, int w, int h, int stride, int qp, orc_itree *t) {
    iplane_t P = {src, stride, w, h, qp, orc_lambda(qp, 1.0), t};
    memset(t->cmode, 0, (size_t)(w / 4) * (h / 4));
    for (int y = 0; y < h; y += ICTU)
        for (int x = 0; x < w; x += ICTU) inode(&P, x, y, 0);
}

/* ---- serialising the decided tree */

typedef struct { uint8_t *b[IS_NS]; size_t n[IS_NS], cap; int over; } isink_t;
static void iput(isink_t *s, int k, unsigned v) {
    if (s->n[k] >= s->cap) { s->over = 1; return; }
    s->b[k][s->n[k]++] = (uint8_t)v;
}
static void iput_leb(isink_t *s, int k, uint32_t u) {
    if (s->n[k] + 5 > s->cap) { s->over = 1; return; }
    s->n[k] = leb_write(s->b[k], s->n[k], u);
}
/* count + levels of the 4x4 block at (x, y) of the level plane */
static void iemit4(isink_t *s, const int16_t *lev, int w, int x, int y) {
    int16_t l[16];
    for (int i = 0; i < 16; ++i) l[i] = lev[(size_t)(y + (i >> 2)) * w + x + (i & 3)];
    int last = -1;
    for (int k = 0; k < 16; ++k) if (l[SCAN4[k]]) last = k;
    iput(s, IS_CNT, (unsigned)(last + 1));
    for (int k = 0; k <= last; ++k) iput_leb(s, IS_LVL, zmap(l[SCAN4[k]]));
}
static void iser(const orc_itree *t, int w, int h, int bx, int by, int d, isink_t *s) {
    const int n = ICTU >> d, hn = n / 2;
    if (n == 4) {
        const int cx = bx / 4, cy = by / 4, mw = w / 4;
        const int L = cx > 0 ? t->cmode[(size_t)cy * mw + cx - 1] : 0;
        const int A = cy > 0 ? t->cmode[(size_t)(cy - 1) * mw + cx] : 0;
        iput(s, L == A ? IS_MODE0 : IS_MODE1, t->cmode[(size_t)cy * mw + cx]);   /* intra_mode_ctx, fdv.h:4194 */
        iemit4(s, t->lev, w, bx, by);
        return;
    }
    const size_t ni = (size_t)(by / n) * inodes_w(w, d) + bx / n;
    const int sp = t->split[d][ni];
    if (sp != 2) iput(s, IS_SPLIT + d, (unsigned)sp);
    if (sp) {
        for (int k = 0; k < 4; ++k) {
            int cx = bx + (k & 1) * hn, cy = by + (k >> 1) * hn;
            if (cx < w && cy < h) iser(t, w, h, cx, cy, d + 1, s);
        }
        return;
    }
    iput(s, IS_MODEB, t->nmode[d][ni]);
    const int rw = (w + 7) / 8;
    for (int ry = by; ry < by + n; ry += 8)
        for (int rx = bx; rx < bx + n; rx += 8) {
            const int f = t->rflag[(size_t)(ry / 8) * rw + rx / 8];
            iput(s, IS_FLAG, (unsigned)f);
            if (f == 1) {
                for (int q = 0; q < 4; ++q) iemit4(s, t->lev, w, rx + (q & 1) * 4, ry + (q >> 1) * 4);
            } else if (f == 2) {
                int16_t l[64];
                for (int i = 0; i < 64; ++i) l[i] = t->lev[(size_t)(ry + (i >> 3)) * w + rx + (i & 7)];
                int last = -1;
                for (int k = 0; k < 64; ++k) if (l[SCAN8[k]]) last = k;
                iput(s, IS_CNT8, (unsigned)(last + 1));
                for (int k = 0; k <= last; ++k) iput_leb(s, IS_LVL8, zmap(l[SCAN8[k]]));
            }
        }
}

/* The container of one plane (fdv.h:4926-5030): the twelve streams, then the
 * smallest of one shared static table, a static table per non-empty stream,
 * and the adaptive coder from flat models (ties to the lower mode). */
size_t orc_image_pack(const orc_itree *t, int w, int h, int qp, uint8_t *out, size_t cap) {
    isink_t s;
    memset(&s, 0, sizeof s);
    s.cap = (size_t)w * h * 5 + 1024;
    for (int k = 0; k < IS_NS; ++k) s.b[k] = malloc(s.cap);
    size_t o = 0;
    uint8_t *all = NULL, *p1 = NULL, *pay[IS_NS] = {0}, *pada = NULL;
    for (int k = 0; k < IS_NS; ++k) if (!s.b[k]) goto done;
    for (int y = 0; y < h; y += ICTU)
        for (int x = 0; x < w; x += ICTU) iser(t, w, h, x, y, 0, &s);
    if (s.over) goto done;
    size_t total = 0;
    for (int k = 0; k < IS_NS; ++k) total += s.n[k];
    all = malloc(total ? total : 1);
    p1 = malloc(total * 2 + 1024);
    if (!all || !p1) goto done;
    uint32_t c1[256] = {0};
    {
        size_t q = 0;
        for (int k = 0; k < IS_NS; ++k) { memcpy(all + q, s.b[k], s.n[k]); q += s.n[k]; }
    }
    for (size_t i = 0; i < total; ++i) c1[all[i]]++;
    uint16_t f1[256], fk[IS_NS][256];
    freq_norm(c1, f1);
    const size_t l1 = rans_enc(all, total, f1, p1, total * 2 + 1024);
    size_t len[IS_NS], hdr = 0;
    for (int k = 0; k < IS_NS; ++k) {
        uint32_t ck[256] = {0};
        for (size_t i = 0; i < s.n[k]; ++i) ck[s.b[k][i]]++;
        freq_norm(ck, fk[k]);
        pay[k] = malloc(s.n[k] * 2 + 1024);
        if (!pay[k]) goto done;
        len[k] = rans_enc(s.b[k], s.n[k], fk[k], pay[k], s.n[k] * 2 + 1024);
        hdr += (size_t)leb_len((uint32_t)s.n[k]);
    }
    const size_t size_one = hdr + freqs_size(f1) + (size_t)leb_len((uint32_t)l1) + l1;
    size_t size_two = hdr;
    for (int k = 0; k < IS_NS; ++k)
        if (s.n[k]) size_two += freqs_size(fk[k]) + (size_t)leb_len((uint32_t)len[k]) + len[k];
    size_t size_ada = (size_t)-1, lada = 0;
    pada = malloc(total * 3 + 4096);
    if (pada) {
        rc_t e = {pada, total * 3 + 4096, 0, 0, 0xFFFFFFFFu, 0, 1, 0};
        const uint32_t flat[256] = {0};
        for (int k = 0; k < IS_NS; ++k) {
            amod_t m; amod_init(&m, flat);
            for (size_t i = 0; i < s.n[k]; ++i) rc_put(&e, &m, s.b[k][i]);
        }
        lada = rc_done(&e);
        if (lada) size_ada = hdr + (size_t)leb_len((uint32_t)lada) + lada;
    }
    int mode = 0; size_t best = size_one;
    if (size_two < best) { best = size_two; mode = 1; }
    if (size_ada < best) { best = size_ada; mode = 2; }
    if (5 + best > cap) goto done;
    out[o++] = (uint8_t)w; out[o++] = (uint8_t)(w >> 8);
    out[o++] = (uint8_t)h; out[o++] = (uint8_t)(h >> 8);
    out[o++] = (uint8_t)qp;
    out[o++] = (uint8_t)mode;
    for (int k = 0; k < IS_NS; ++k) o = leb_write(out, o, (uint32_t)s.n[k]);
    if (mode == 2) {
        o = leb_write(out, o, (uint32_t)lada); memcpy(out + o, pada, lada); o += lada;
    } else if (mode == 0) {
        o = freqs_write(out, o, f1); o = leb_write(out, o, (uint32_t)l1); memcpy(out + o, p1, l1); o += l1;
    } else {
        for (int k = 0; k < IS_NS; ++k) {
            if (!s.n[k]) continue;
            o = freqs_write(out, o, fk[k]); o = leb_write(out, o, (uint32_t)len[k]);
            memcpy(out + o, pay[k], len[k]); o += len[k];
        }
    }
done:
    for (int k = 0; k < IS_NS; ++k) { free(s.b[k]); free(pay[k]); }
    free(all); free(p1); free(pada);
    return o;
}

int orc_itree_alloc(orc_itree *t, int w, int h) {
    memset(t, 0, sizeof *t);
    t->rec = malloc((size_t)w * h);
    t->lev = calloc((size_t)w * h, sizeof *t->lev);
    t->rflag = calloc((size_t)((w + 7) / 8) * ((h + 7) / 8), 1);
    t->cmode = calloc((size_t)(w / 4) * (h / 4), 1);
    int ok = t->rec && t->lev && t->rflag && t->cmode;
    for (int d = 0; d < 4; ++d) {
        size_t nn = (size_t)inodes_w(w, d) * inodes_w(h, d);
        t->split[d] = calloc(nn, 1);
        t->nmode[d] = calloc(nn, 1);
        ok = ok && t->split[d] && t->nmode[d];
    }
    if (!ok) { orc_itree_free(t); return -1; }
    return 0;
}
void orc_itree_free(orc_itree *t) {
    free(t->rec); free(t->lev); free(t->rflag); free(t->cmode);
    for (int d = 0; d < 4; ++d) { free(t->split[d]); free(t->nmode[d]); }
    memset(t, 0, sizeof *t);
}

size_t orc_image_encode(const uint8_t *src, int w, int h, int stride, int qp, uint8_t *out, size_t cap,
                        uint8_t *recon) {
    if (w <= 0 || h <= 0 || (w & 3) || (h & 3)) return 0;
    orc_itree t;
    if (orc_itree_alloc(&t, w, h) != 0) return 0;
    orc_image_analyse(src, w, h, stride, qp, &t);
    orc_deblock_plane(t.rec, w, h, w, qp);
    if (recon) memcpy(recon, t.rec, (size_t)w * h);
    size_t n = orc_image_pack(&t, w, h, qp, out, cap);
    orc_itree_free(&t);
    return n;
}

/* A key frame: the three planes coded separately, each behind a 32-bit
 * little-endian length (iframe_encode, fdv.h:5487-5539). */
size_t orc_iframe_encode(const uint8_t *y, const uint8_t *u, const uint8_t *v, int w, int h, int qp,
                         uint8_t *out, size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv) {
    const uint8_t *src[3] = {y, u, v};
    uint8_t *rec[3] = {ry, ru, rv};
    const int pw[3] = {w, w / 2, w / 2}, ph[3] = {h, h / 2, h / 2};
    size_t o = 0;
    for (int i = 0; i < 3; ++i) {
        size_t bcap = (size_t)pw[i] * ph[i] * 2 + 4096;
        uint8_t *blob = malloc(bcap);
        if (!blob) return 0;
        size_t n = orc_image_encode(src[i], pw[i], ph[i], pw[i], qp, blob, bcap, rec[i]);
        if (!n || o + 4 + n > cap) { free(blob); return 0; }
        out[o++] = (uint8_t)n; out[o++] = (uint8_t)(n >> 8); out[o++] = (uint8_t)(n >> 16); out[o++] = (uint8_t)(n >> 24);
        memcpy(out + o, blob, n);
        o += n;
        free(blob);
    }
    return o;
}

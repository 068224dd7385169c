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
// Inside an MB the four wavefronts first run the independent searches side by
// side -- the 16x16 diamond on reference 0, on reference 1, the four chained
// 8x8 quadrant diamonds, and the intra SATD screen -- each wave-local (v_sad_u8
// on packed bytes, DPP reductions, scalar control).  The reference evaluates
// some of these only when earlier candidates leave room (fdv.h:5913, 5950,
// 5996); the searches are pure functions of the MB and its neighbours, so
// computing them early changes no result, only latency.  Residual coding then
// takes the four wavefronts as the four 8x8 regions (fdvgpu_resid.h).
//
// All rate-distortion costs are IEEE doubles computed without contraction
// (-ffp-contract=off) in the reference's operation order.
#include <hip/hip_runtime.h>
#include "fdvgpu_common.h"
#include "fdvgpu_dev.h"
#include "fdvgpu_resid.h"

// Sub-pel refinement reads four planes around the whole-pel best vector: G
// (integer samples), b and h (horizontal / vertical half samples) and j (centre),
// for block-relative X, Y in [-1, bs] (stride SPN).  Built once per search from
// an unclipped horizontal-tap plane, they turn every quarter-pel sample into the
// rounded mean of two lookups (H.264 fig. 8-4; fdv.h:2931-2990).
#define SPN 18
struct SubpelWs {
    uint8_t pl[4 * SPN * SPN];
    int16_t hr[SPN * (16 + 7)];
};
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
    RegionWs ws;
    SubpelWs sp[3];                   // half-pel planes of the three searching waves
    int32_t red[2][4][4];
    int32_t me[8][4];                 // [0],[1]: 16x16 on ref 0/1 (mvx, mvy, sad); [2..5]:
                                      // quadrants; [6]: intra SATD per mode; [7][0]: quadrant SAD sum
    uint64_t hand[3][FDVG_HWORDS];
    uint8_t hbytes[48];
    uint8_t lcol[32];                 // previous MB's right recon column: Y 16, U 8, V 8
    int32_t row;
};

// ---- wave-local motion search ------------------------------------------------

__device__ __forceinline__ uint32_t win_word(const uint8_t *w, int y, int x) {
    // 4 bytes at an arbitrary x of window row y, from two aligned words
    const uint32_t *r = (const uint32_t *)(w + y * WIN);
    const int a = x >> 2;
    return __builtin_amdgcn_alignbyte(r[a + 1], r[a], (unsigned)(x & 3));
}

// SADs of up to 8 whole-pel vectors (quarter-pel units, multiples of 4) for the
// bs x bs block at (ox, oy) of the MB: 8 lanes per point.  Results are uniform.
__device__ __forceinline__ void pts_int8(const uint8_t *w, const uint8_t *cur, int bs, int ox, int oy,
                                         const int *px, const int *py, int *sad) {
    const int L = threadIdx.x & 63, k = L >> 3, j = L & 7;
    int mx = 0, my = 0;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) if (kk == k) { mx = px[kk]; my = py[kk]; }
    const int dx = mx >> 2, dy = my >> 2;
    uint32_t acc = 0;
    if (bs == 16) {
#pragma unroll
        for (int yy = 0; yy < 2; ++yy) {
            const int y = 2 * j + yy;
            const uint32_t *cw = (const uint32_t *)(cur + y * 16);
            const int Y = WOFF + y + dy, X = WOFF + dx;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = __builtin_amdgcn_sad_u8(cw[q], win_word(w, Y, X + 4 * q), acc);
        }
    } else {
        const uint32_t *cw = (const uint32_t *)(cur + (oy + j) * 16 + ox);
        const int Y = WOFF + oy + j + dy, X = WOFF + ox + dx;
        acc = __builtin_amdgcn_sad_u8(cw[0], win_word(w, Y, X), acc);
        acc = __builtin_amdgcn_sad_u8(cw[1], win_word(w, Y, X + 4), acc);
    }
    const int t = oct_sum8((int)acc);
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) sad[kk] = __builtin_amdgcn_readlane(t, 8 * kk + 4);
}

// quarter-pel phase (fy*4+fx) -> two (plane, dx, dy) sources
__device__ __constant__ int8_t k_PHSRC[16][6] = {
    {0, 0, 0, 0, 0, 0}, {0, 0, 0, 1, 0, 0}, {1, 0, 0, 1, 0, 0}, {1, 0, 0, 0, 1, 0},
    {0, 0, 0, 2, 0, 0}, {1, 0, 0, 2, 0, 0}, {1, 0, 0, 3, 0, 0}, {1, 0, 0, 2, 1, 0},
    {2, 0, 0, 2, 0, 0}, {2, 0, 0, 3, 0, 0}, {3, 0, 0, 3, 0, 0}, {3, 0, 0, 2, 1, 0},
    {2, 0, 0, 0, 0, 1}, {2, 0, 0, 1, 0, 1}, {3, 0, 0, 1, 0, 1}, {2, 1, 0, 1, 0, 1}};

__device__ __forceinline__ void subpel_planes(const uint8_t *w, SubpelWs &P, int bs, int ox, int oy,
                                              int dx, int dy) {
    const int L = threadIdx.x & 63, n = bs + 2, X0 = WOFF + ox + dx - 1, Y0 = WOFF + oy + dy;
    for (int i = L; i < n * (bs + 7); i += 64) {               // hraw rows Y = -3 .. bs+3
        const int X = i % n, Y = i / n;
        P.hr[Y * SPN + X] = (int16_t)whr(w, X0 + X, Y0 - 3 + Y);
    }
    wave_sync();
    for (int i = L; i < n * n; i += 64) {                      // X, Y = -1 .. bs
        const int X = i % n, Y = i / n, o = Y * SPN + X;
        const int Wx = X0 + X, Wy = Y0 - 1 + Y;
        P.pl[o] = w[Wy * WIN + Wx];
        P.pl[SPN * SPN + o] = (uint8_t)clip8((P.hr[(Y + 2) * SPN + X] + 16) >> 5);
        P.pl[2 * SPN * SPN + o] = (uint8_t)clip8((wvr(w, Wx, Wy) + 16) >> 5);
        const int jj = tap6(P.hr[Y * SPN + X], P.hr[(Y + 1) * SPN + X], P.hr[(Y + 2) * SPN + X],
                            P.hr[(Y + 3) * SPN + X], P.hr[(Y + 4) * SPN + X], P.hr[(Y + 5) * SPN + X]);
        P.pl[3 * SPN * SPN + o] = (uint8_t)clip8((jj + 512) >> 10);
    }
    wave_sync();
}

// SADs of 4 quarter-pel vectors within 3 quarter pels of the whole-pel vector
// (ix, iy) the planes were built for: 16 lanes per point, no divergence.
__device__ __forceinline__ void pts_frac4(const SubpelWs &P, const uint8_t *cur, int bs, int ox, int oy,
                                          int ix, int iy, const int *px, const int *py, int *sad) {
    const int L = threadIdx.x & 63, k = L >> 4, j = L & 15;
    int mx = 0, my = 0;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) if (kk == k) { mx = px[kk]; my = py[kk]; }
    const int dxq = mx - ix, dyq = my - iy;
    const int ph = ((dyq & 3) << 2) | (dxq & 3), bxo = (dxq >> 2) + 1, byo = (dyq >> 2) + 1;
    const int a0 = k_PHSRC[ph][0] * SPN * SPN + (byo + k_PHSRC[ph][2]) * SPN + bxo + k_PHSRC[ph][1];
    const int a1 = k_PHSRC[ph][3] * SPN * SPN + (byo + k_PHSRC[ph][5]) * SPN + bxo + k_PHSRC[ph][4];
    int acc = 0;
    if (bs == 16) {
#pragma unroll
        for (int x = 0; x < 16; ++x) {
            const int v = (P.pl[a0 + j * SPN + x] + P.pl[a1 + j * SPN + x] + 1) >> 1;
            const int d = cur[j * 16 + x] - v;
            acc += d < 0 ? -d : d;
        }
    } else {
        const int y = j >> 1, x0 = (j & 1) * 4;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const int v = (P.pl[a0 + y * SPN + x0 + x] + P.pl[a1 + y * SPN + x0 + x] + 1) >> 1;
            const int d = cur[(oy + y) * 16 + ox + x0 + x] - v;
            acc += d < 0 ? -d : d;
        }
    }
    acc = row_sum16(acc);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) sad[kk] = __builtin_amdgcn_readlane(acc, 16 * kk);
}

// The reference's EPZS diamond + sub-pel cross (fdv_me_search_inner,
// fdv.h:3375-3466) on one wavefront: (0,0) and the whole-pel-rounded, clamped
// predictor; large diamonds (8 points, strict improvement, first wins, at most
// 64 rounds), one small diamond, then the half- and quarter-pel crosses
// (unclamped, skipped at SAD 0).  Returns the SAD, vector in (*omx, *omy).
__device__ __noinline__ int me_diamond(const uint8_t *w, const uint8_t *cur, SubpelWs &SP, int bs, int ox, int oy,
                          int pmx, int pmy, int *omx, int *omy) {
    const int lim = FDVG_LIM;
    int px = (pmx >> 2) * 4, py = (pmy >> 2) * 4;
    px = px < -lim ? -lim : px > lim ? lim : px;
    py = py < -lim ? -lim : py > lim ? lim : py;
    int P[8], Q[8], S[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { P[k] = 0; Q[k] = 0; }
    P[1] = px; Q[1] = py;
    pts_int8(w, cur, bs, ox, oy, P, Q, S);
    int best = S[0], bxv = 0, byv = 0;
    if ((px || py) && S[1] < best) { best = S[1]; bxv = px; byv = py; }
    for (int it = 0; it < 64; ++it) {
        const int cx = bxv, cy = byv;
        int valid = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int nx = cx + k_LDSP[k][0], ny = cy + k_LDSP[k][1];
            const bool ok = !(nx < -lim || nx > lim || ny < -lim || ny > lim);
            valid |= (int)ok << k;
            P[k] = ok ? nx : cx; Q[k] = ok ? ny : cy;
        }
        pts_int8(w, cur, bs, ox, oy, P, Q, S);
        int moved = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (((valid >> k) & 1) && S[k] < best) { best = S[k]; bxv = P[k]; byv = Q[k]; moved = 1; }
        if (!moved) break;
    }
    {
        const int cx = bxv, cy = byv;
        int valid = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int nx = k < 4 ? cx + k_SDSP[k & 3][0] : cx, ny = k < 4 ? cy + k_SDSP[k & 3][1] : cy;
            const bool ok = k < 4 && !(nx < -lim || nx > lim || ny < -lim || ny > lim);
            valid |= (int)ok << k;
            P[k] = ok ? nx : cx; Q[k] = ok ? ny : cy;
        }
        pts_int8(w, cur, bs, ox, oy, P, Q, S);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (((valid >> k) & 1) && S[k] < best) { best = S[k]; bxv = P[k]; byv = Q[k]; }
    }
    const int ix = bxv, iy = byv;
    if (best > 0) subpel_planes(w, SP, bs, ox, oy, ix >> 2, iy >> 2);
    for (int st = 2; st >= 1 && best > 0; st >>= 1) {
        const int cx = bxv, cy = byv;
#pragma unroll
        for (int k = 0; k < 4; ++k) { P[k] = cx + k_CROSS[k][0] * st; Q[k] = cy + k_CROSS[k][1] * st; }
        pts_frac4(SP, cur, bs, ox, oy, ix, iy, P, Q, S);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (S[k] < best) { best = S[k]; bxv = P[k]; byv = Q[k]; }
    }
    *omx = bxv; *omy = byv;
    return best;
}

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
__device__ __noinline__ void satd_modes(const uint8_t *cur, const Nb &nb, const uint8_t *top,
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

// ---- residual coding of a whole MB's luma (four regions, one per wavefront)
//      or of the two chroma blocks (waves 0, 1)
__device__ __forceinline__ void code_luma(Smem &s, const uint8_t *pred, uint8_t *rec, int16_t *lev,
                                          uint8_t *flag, int qp, double lambda) {
    const int w = threadIdx.x >> 6, o = (w >> 1) * 128 + (w & 1) * 8;
    region_code<0>(s.ws, w, s.cur + o, 16, pred + o, 16, rec + o, 16, lev, flag, qp, lambda);
    __syncthreads();
}
template <int KIND>
__device__ __forceinline__ void code_chroma(Smem &s, int qp, double lambda) {
    const int w = threadIdx.x >> 6;
    if (w < 2)
        region_code<KIND>(s.ws, w, s.ccur + 64 * w, 8, s.cpred + 64 * w, 8, s.crec + 64 * w, 8,
                          s.clev, s.cflag, qp, lambda);
    __syncthreads();
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

// Optional per-phase cycle accounting (thread 0 of every workgroup), compiled in
// only with -DFDVG_PHASE_PROF for performance investigation.
#ifdef FDVG_PHASE_PROF
__device__ unsigned long long g_phase[16];
#define PH(k) do { if (tid == 0) { long long t_ = clock64(); ph[k] += t_ - ph_last; ph_last = t_; } } while (0)
extern "C" int fdvg_phase_read(unsigned long long *out) {
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase), sizeof g_phase) == hipSuccess ? 0 : -1;
}
#else
#define PH(k) do { } while (0)
#endif

extern "C" __global__ void __launch_bounds__(256)
k_analyse(const fdvg_seq *__restrict__ seqs, const int2 *__restrict__ rowtab, int nrows,
          unsigned *ticket, fdvg_mbrec *__restrict__ recs, uint64_t *__restrict__ hand) {
    __shared__ Smem s;
    const int tid = threadIdx.x, wave = tid >> 6;
    int par = 0;
#ifdef FDVG_PHASE_PROF
    long long ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long ph_last = clock64();
#endif
    for (;;) {
        if (tid == 0) s.row = (int)atomicAdd(ticket, 1u);
        __syncthreads();
        const int row = s.row;
        __syncthreads();
        if (row >= nrows) {
#ifdef FDVG_PHASE_PROF
            if (tid == 0) for (int k = 0; k < 8; ++k) atomicAdd(&g_phase[k], (unsigned long long)ph[k]);
#endif
            return;
        }
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
            PH(0);

            // ---- neighbours from the row above / the previous MB
            const int ht = my > 0, hl = mx > 0;
            if (tid < 16) s.hbytes[1 + tid] = ht ? (uint8_t)hbyte(s, 1, tid) : 0;
            if (tid == 16) s.hbytes[0] = (ht && hl) ? (uint8_t)hbyte(s, 0, 15) : 128;
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
            __syncthreads();

            const fdvg_mbinfo *gv = q.given ? &q.given[mb - q.mb0] : nullptr;
            const fdvg_metab *mt = (q.me_mode == FDVG_ME_TABLE && q.metab) ? &q.metab[mb - q.mb0] : nullptr;
            int mode, bmi = 0, iref = 0, imx = 0, imy = 0, nsub = 0;
            int qmv[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
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
                {
                    const int e0 = wave_sum(e[0]), e1 = wave_sum(e[1]), e2 = wave_sum(e[2]);
                    if ((tid & 63) == 0) { s.red[par][wave][0] = e0; s.red[par][wave][1] = e1; s.red[par][wave][2] = e2; }
                }
                __syncthreads();
                int Ds = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (c >= ncand) continue;
                    const int ec = s.red[par][0][c] + s.red[par][1][c] + s.red[par][2][c] + s.red[par][3][c];
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
                PH(1);

                // ---- the independent searches, one per wavefront
                const Nb nbL = nb_setup(&s.hbytes[1], s.lcol, s.hbytes[0], 16, ht, hl);
                if (!gv && !skip_wins) {
                    if (wave == 0 || (wave == 1 && q.navail == 2)) {
                        int a, b, sd;
                        if (mt) { a = mt->mv16[wave][0]; b = mt->mv16[wave][1]; sd = mt->sad16[wave]; }
                        else sd = me_diamond(s.win[wave], s.cur, s.sp[wave], 16, 0, 0, pmx, pmy, &a, &b);
                        if ((tid & 63) == 0) { s.me[wave][0] = a; s.me[wave][1] = b; s.me[wave][2] = sd; }
                    } else if (wave == 2) {
                        int sx = pmx, sy = pmy, ssad = 0;
                        for (int qd = 0; qd < 4; ++qd) {
                            int a, b;
                            if (mt) { a = mt->mv8[qd][0]; b = mt->mv8[qd][1]; ssad += mt->sad8[qd]; }
                            else ssad += me_diamond(s.win[0], s.cur, s.sp[2], 8, (qd & 1) * 8, (qd >> 1) * 8, sx, sy, &a, &b);
                            if ((tid & 63) == 0) { s.me[2 + qd][0] = a; s.me[2 + qd][1] = b; }
                            sx = a; sy = b;
                        }
                        if ((tid & 63) == 0) s.me[7][0] = ssad;
                    } else if (wave == 3) {
                        int sat[4];
                        satd_modes(s.cur, nbL, &s.hbytes[1], s.lcol, sat);
                        if ((tid & 63) == 0) for (int m = 0; m < 4; ++m) s.me[6][m] = sat[m];
                    }
                }
                __syncthreads();
                PH(2);

                // ---- INTER16 over the references (fdv.h:5909-5943)
                double Ji = 1e30;
                int sad16 = 0;
                const bool want_i = gv ? gv->mode == 1 : !skip_wins;
                for (int r = 0; r < q.navail && want_i; ++r) {
                    if (gv && r != gv->ref) continue;
                    if (!gv && r > 0 && Jbest <= lambda * (double)FL_INTER16) break;
                    int rmx, rmy, rsad = 0;
                    if (gv) { rmx = gv->mv[0][0]; rmy = gv->mv[0][1]; }
                    else { rmx = s.me[r][0]; rmy = s.me[r][1]; rsad = s.me[r][2]; }
                    s.pred[tid] = (uint8_t)lumaq(s.win[r], tid & 15, tid >> 4, rmx, rmy);
                    __syncthreads();
                    code_luma(s, s.pred, s.recT, s.levT, s.flagT, qp, lambda);
                    const int bb = 6 + egbits(zmap(rmx - pmx)) + egbits(zmap(rmy - pmy)) +
                                   s.ws.b[0] + s.ws.b[1] + s.ws.b[2] + s.ws.b[3];
                    const double dd = (double)(s.ws.e[0] + s.ws.e[1] + s.ws.e[2] + s.ws.e[3]);
                    const double J = dd + lambda * (double)bb;
                    if (J < Jbest) Jbest = J;
                    if (J < Ji) {
                        Ji = J; iref = r; imx = rmx; imy = rmy; sad16 = rsad;
                        s.recI[tid] = s.recT[tid];
                        s.levI[tid] = s.levT[tid];
                        if (tid < 4) s.flagI[tid] = s.flagT[tid];
                    }
                    __syncthreads();
                    PH(3);
                }

                // ---- INTRA 16x16: SATD screen, one coding (fdv.h:5950-5988)
                double Jn = 1e30;
                const bool want_n = gv ? gv->mode == 2 : (!skip_wins && Jbest > lambda * (double)FL_INTRA);
                if (want_n) {
                    int sel = 0;
                    if (gv) sel = gv->sub;
                    else {
                        int sbest = 0; sel = -1;
                        for (int m = 0; m < 4; ++m) {
                            if (!nn_ok(m, nbL)) continue;
                            if (sel < 0 || s.me[6][m] < sbest) { sbest = s.me[6][m]; sel = m; }
                        }
                    }
                    s.pred[tid] = (uint8_t)ipred(sel, nbL, &s.hbytes[1], s.lcol, 16, tid & 15, tid >> 4);
                    __syncthreads();
                    code_luma(s, s.pred, s.recN, s.levN, s.flagN, qp, lambda);
                    const int bn = 6 + s.ws.b[0] + s.ws.b[1] + s.ws.b[2] + s.ws.b[3];
                    Jn = (double)(s.ws.e[0] + s.ws.e[1] + s.ws.e[2] + s.ws.e[3]) + lambda * (double)bn;
                    nsub = sel;
                    if (Jn < Jbest) Jbest = Jn;
                }
                PH(4);

                // ---- INTER8: the quadrant vectors, the SAD split gate, four 8x8
                //      residuals (fdv.h:5996-6033)
                double J8 = 1e30;
                bool want_8 = gv ? gv->mode == 3 : (!skip_wins && Jbest > lambda * (double)FL_INTER8);
                if (want_8) {
#pragma unroll
                    for (int qd = 0; qd < 4; ++qd) {
                        qmv[qd][0] = gv ? gv->mv[qd][0] : s.me[2 + qd][0];
                        qmv[qd][1] = gv ? gv->mv[qd][1] : s.me[2 + qd][1];
                    }
                    const int ssad = s.me[7][0];
                    PH(5);
                    if (!gv && sad16 > 0 && ssad * 10 >= sad16 * 9) want_8 = false;
                    if (want_8) {
                        const int x = tid & 15, y = tid >> 4, qd = (y >> 3) * 2 + (x >> 3);
                        const int vx = qd == 0 ? qmv[0][0] : qd == 1 ? qmv[1][0] : qd == 2 ? qmv[2][0] : qmv[3][0];
                        const int vy = qd == 0 ? qmv[0][1] : qd == 1 ? qmv[1][1] : qd == 2 ? qmv[2][1] : qmv[3][1];
                        s.pred[tid] = (uint8_t)lumaq(s.win[0], x, y, vx, vy);
                        __syncthreads();
                        code_luma(s, s.pred, s.rec8, s.lev8, s.flag8, qp, lambda);
                        int b8 = 3 + s.ws.b[0] + s.ws.b[1] + s.ws.b[2] + s.ws.b[3];
                        int qpx = pmx, qpy = pmy;
#pragma unroll
                        for (int qd2 = 0; qd2 < 4; ++qd2) {
                            b8 += egbits(zmap(qmv[qd2][0] - qpx)) + egbits(zmap(qmv[qd2][1] - qpy));
                            qpx = qmv[qd2][0]; qpy = qmv[qd2][1];
                        }
                        J8 = (double)(s.ws.e[0] + s.ws.e[1] + s.ws.e[2] + s.ws.e[3]) + lambda * (double)b8;
                    }
                    PH(6);
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
                code_chroma<1>(s, qp, lambda);
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
                code_chroma<0>(s, qp, lambda);
                ckind = 2;
            } else {
                if (tid < 128) {
                    const int pl = tid >> 6, p = tid & 63, x = p & 7, y = p >> 3, qd = (y >> 2) * 2 + (x >> 2);
                    const int vx = qd == 0 ? qmv[0][0] : qd == 1 ? qmv[1][0] : qd == 2 ? qmv[2][0] : qmv[3][0];
                    const int vy = qd == 0 ? qmv[0][1] : qd == 1 ? qmv[1][1] : qd == 2 ? qmv[2][1] : qmv[3][1];
                    s.cpred[tid] = (uint8_t)chromaq(s.cwin[0][pl], x, y, vx / 2, vy / 2);
                }
                __syncthreads();
                code_chroma<2>(s, qp, lambda);
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
            PH(7);
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

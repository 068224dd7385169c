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
#include "fdvgpu_mb.h"

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
    uint8_t lcol[32];                 // left neighbour's right recon column: Y 16, U 8, V 8
    uint8_t ctop[2][9];               // chroma: corner, then the 8 samples above (U, V)
    int32_t row;
};

// Inputs of one MB's decision that come from its neighbours: the running
// vector predictor, the merge candidates (fdv.h:5711-5732) and whether the
// top / left neighbours exist.  Luma intra neighbours are in s.hbytes (corner,
// then 16 above) and s.lcol; chroma ones in s.ctop and s.lcol.
// The small arrays are vector values (elements 0..2 / x,y pairs): indexing
// them with a run-time value then selects among registers, where an array
// member would put the whole struct in scratch memory.
typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i8v __attribute__((ext_vector_type(8)));
struct MbCtx {
    int pmx, pmy, ncand, ht, hl;
    i4v cx, cy, cr;
};
// The decision and what the finishing stage needs of it; qmv[2 qd + 0/1] is
// the vector of quadrant qd.
struct MbDec {
    int mode, bmi, iref, imx, imy, nsub;
    i8v qmv;
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

// ---- the MB body, shared by the wavefront kernel and the block-parallel ones

// Stage the source MB and, when the MB may use them, the 64x64 luma and 32x32
// chroma windows of each reference around it (of the references in refmask,
// for a caller that knows which the MB's decision reads).  The caller
// synchronises.
__device__ __forceinline__ void mb_stage(Smem &s, const fdvg_seq &q, int mb, int bx, int by, int gmode = -1,
                                         int refmask = 3) {
    const int tid = threadIdx.x, cw = q.w >> 1, cbx = bx >> 1, cby = by >> 1;
    s.cur[tid] = gmem(q.cy)[(size_t)(by + (tid >> 4)) * q.w + bx + (tid & 15)];
    if (tid < 128) {
        const int pl = tid >> 6, p = tid & 63;
        const gmem_t<uint8_t> *src = gmem(pl ? q.cv : q.cu);
        s.ccur[tid] = src[(size_t)(cby + (p >> 3)) * cw + cbx + (p & 7)];
    }
    // gmode >= 0: the MB's given mode, passed by a caller that holds it
    const bool intra = gmode >= 0 ? gmode == 2 : (q.given && q.given[mb - q.mb0].mode == 2);
    const bool need_ref = !q.force_skip && !intra;
    const int nwin = q.force_skip ? 1 : q.navail;
    if (need_ref || q.force_skip) {
        for (int r = 0; r < nwin; ++r) {
            if (!((refmask >> r) & 1)) continue;     // uniform: a reference the caller does not use
            // 64 rows x 64 bytes: 16 bytes a thread, two 8-byte loads (8-aligned)
            const int wy = tid >> 2, wx = (tid & 3) * 16;
            const gmem_t<uint8_t> *src = gmem(r ? q.ry[1] : q.ry[0]) + (ptrdiff_t)(by - WOFF + wy) * q.ys + bx - WOFF + wx;
            uint64_t a = *(const gmem_t<uint64_t> *)src, b = *(const gmem_t<uint64_t> *)(src + 8);
            *(uint64_t *)&s.win[r][wy * WIN + wx] = a;
            *(uint64_t *)&s.win[r][wy * WIN + wx + 8] = b;
            // 2 planes x 32 rows x 32 bytes: 8 bytes a thread (4-aligned)
            const int pl = tid >> 7, t2 = tid & 127, cyy = t2 >> 2, cxx = (t2 & 3) * 8;
            const gmem_t<uint8_t> *cp = gmem(pl ? (r ? q.rv[1] : q.rv[0]) : (r ? q.ru[1] : q.ru[0])) +
                                        (ptrdiff_t)(cby - CWOFF + cyy) * q.cs + cbx - CWOFF + cxx;
            const gmem_t<uint32_t> *c32 = (const gmem_t<uint32_t> *)cp;
            uint32_t u0 = c32[0], u1 = c32[1];
            *(uint32_t *)&s.cwin[r][pl][cyy * CWIN + cxx] = u0;
            *(uint32_t *)&s.cwin[r][pl][cyy * CWIN + cxx + 4] = u1;
        }
    }
}

// The mode decision of one MB (pframe_encode's MB loop, fdv.h:5820-6040): SKIP
// over the merge candidates, INTER16 per reference, INTRA 16x16, INTER8, then
// the strict-< commit.  With `gv` the decision is given and only its mode's
// candidate is built; with `mt` the motion vectors come from a table instead of
// the diamond search.  Leaves the chosen luma reconstruction / levels / flags in
// s.recS (SKIP), s.recI/levI/flagI (INTER16), s.recN/... (INTRA), s.rec8/...
// (INTER8).
__device__ __forceinline__ void mb_decide(Smem &s, const fdvg_seq &q, MbCtx &c, const fdvg_mbinfo *gv,
                                          const fdvg_metab *mt, int &par, MbDec &D) {
    const int tid = threadIdx.x, wave = tid >> 6;
    const double lambda = q.lambda;
    const int qp = q.qp, pmx = c.pmx, pmy = c.pmy;
    D.bmi = 0; D.iref = 0; D.imx = 0; D.imy = 0; D.nsub = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) { D.qmv[2 * k + 0] = 0; D.qmv[2 * k + 1] = 0; }

    if (q.force_skip) {
        // a held frame: SKIP with the zero vector everywhere (fdv.h:5820-5836)
        D.mode = 0; c.ncand = 1; c.cx[0] = 0; c.cy[0] = 0; c.cr[0] = 0;
        s.recS[tid] = s.win[0][(WOFF + (tid >> 4)) * WIN + WOFF + (tid & 15)];
        __syncthreads();
        return;
    }
    // ---- SKIP: best merge candidate by SSD, first wins (fdv.h:5847-5866)
    const int ncand = c.ncand;
    const int x = tid & 15, y = tid >> 4;
    i4v pv = {0, 0, 0, 0}, e = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= ncand || (gv && !(gv->mode == 0 && gv->merge_idx == k))) continue;
        pv[k] = lumaq(s.win[c.cr[k]], x, y, c.cx[k], c.cy[k]);
        const int d = s.cur[tid] - pv[k];
        e[k] = d * d;
    }
    {
        const int e0 = wave_sum(e[0]), e1 = wave_sum(e[1]), e2 = wave_sum(e[2]);
        if ((tid & 63) == 0) { s.red[par][wave][0] = e0; s.red[par][wave][1] = e1; s.red[par][wave][2] = e2; }
    }
    __syncthreads();
    int Ds = 0, bmi = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= ncand) continue;
        const int ec = s.red[par][0][k] + s.red[par][1][k] + s.red[par][2][k] + s.red[par][3][k];
        if (gv) { if (gv->mode == 0 && gv->merge_idx == k) { Ds = ec; bmi = k; } }
        else if (k == 0 || ec < Ds) { Ds = ec; bmi = k; }
    }
    par ^= 1;
    D.bmi = bmi;
    s.recS[tid] = (uint8_t)(bmi == 0 ? pv[0] : bmi == 1 ? pv[1] : pv[2]);
    const double Js = (double)Ds + lambda * (double)(3 + (ncand > 1 ? 3 : 0));
    double Jbest = Js;
    bool skip_wins = Js <= lambda * (double)FL_INTRA;
    if (gv) skip_wins = false;
    const bool pruned = mt && (mt->flags & 1);       // the proposer expects SKIP: no INTRA, no INTER8

    // ---- the independent searches, one per wavefront
    const Nb nbL = nb_setup(&s.hbytes[1], s.lcol, s.hbytes[0], 16, c.ht, c.hl);
    const bool screen = gv && gv->mode == 2 && gv->sub == FDVG_SUB_SCREEN;
    if ((!gv && !skip_wins) || screen) {
        // a given INTRA MB runs only the sub-mode screen (wave 3)
        if (!screen && (wave == 0 || (wave == 1 && q.navail == 2))) {
            int a, b, sd;
            if (mt) { a = mt->mv16[wave][0]; b = mt->mv16[wave][1]; sd = mt->sad16[wave]; }
            else sd = me_diamond(s.win[wave], s.cur, s.sp[wave], 16, 0, 0, pmx, pmy, &a, &b);
            if ((tid & 63) == 0) { s.me[wave][0] = a; s.me[wave][1] = b; s.me[wave][2] = sd; }
        } else if (!screen && wave == 2) {
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
            Ji = J; D.iref = r; D.imx = rmx; D.imy = rmy; sad16 = rsad;
            s.recI[tid] = s.recT[tid];
            s.levI[tid] = s.levT[tid];
            if (tid < 4) s.flagI[tid] = s.flagT[tid];
        }
        __syncthreads();
    }

    // ---- INTRA 16x16: SATD screen, one coding (fdv.h:5950-5988)
    double Jn = 1e30;
    const bool want_n = gv ? gv->mode == 2 : (!skip_wins && !pruned && Jbest > lambda * (double)FL_INTRA);
    if (want_n) {
        int sel = 0;
        if (gv && !screen) sel = gv->sub;
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
        D.nsub = sel;
        if (Jn < Jbest) Jbest = Jn;
    }

    // ---- INTER8: the quadrant vectors, the SAD split gate, four 8x8 residuals
    //      (fdv.h:5996-6033)
    double J8 = 1e30;
    bool want_8 = gv ? gv->mode == 3 : (!skip_wins && !pruned && Jbest > lambda * (double)FL_INTER8);
    if (want_8) {
#pragma unroll
        for (int qd = 0; qd < 4; ++qd) {
            D.qmv[2 * qd + 0] = gv ? gv->mv[qd][0] : s.me[2 + qd][0];
            D.qmv[2 * qd + 1] = gv ? gv->mv[qd][1] : s.me[2 + qd][1];
        }
        const int ssad = s.me[7][0];
        if (!gv && sad16 > 0 && ssad * 10 >= sad16 * 9) want_8 = false;
        if (want_8) {
            const int qd = (y >> 3) * 2 + (x >> 3);
            const int vx = qd == 0 ? D.qmv[0] : qd == 1 ? D.qmv[2] : qd == 2 ? D.qmv[4] : D.qmv[6];
            const int vy = qd == 0 ? D.qmv[1] : qd == 1 ? D.qmv[3] : qd == 2 ? D.qmv[5] : D.qmv[7];
            s.pred[tid] = (uint8_t)lumaq(s.win[0], x, y, vx, vy);
            __syncthreads();
            code_luma(s, s.pred, s.rec8, s.lev8, s.flag8, qp, lambda);
            int b8 = 3 + s.ws.b[0] + s.ws.b[1] + s.ws.b[2] + s.ws.b[3];
            int qpx = pmx, qpy = pmy;
#pragma unroll
            for (int qd2 = 0; qd2 < 4; ++qd2) {
                b8 += egbits(zmap(D.qmv[2 * qd2 + 0] - qpx)) + egbits(zmap(D.qmv[2 * qd2 + 1] - qpy));
                qpx = D.qmv[2 * qd2 + 0]; qpy = D.qmv[2 * qd2 + 1];
            }
            J8 = (double)(s.ws.e[0] + s.ws.e[1] + s.ws.e[2] + s.ws.e[3]) + lambda * (double)b8;
        }
    }

    // ---- commit: strict < in order SKIP, INTER16, INTRA, INTER8
    if (gv) D.mode = gv->mode;
    else {
        double Jm = Js; D.mode = 0;
        if (Ji < Jm) { Jm = Ji; D.mode = 1; }
        if (Jn < Jm) { Jm = Jn; D.mode = 2; }
        if (J8 < Jm) { Jm = J8; D.mode = 3; }
    }
}

// The motion record an MB leaves for its neighbours (fdv_mbmv, fdv.h:5680):
// vector, reference, whether it is inter.
__device__ __forceinline__ void mb_motion(const MbCtx &c, const MbDec &D, int *nmx, int *nmy, int *nref,
                                          int *ninter) {
    const int b = D.bmi;
    if (D.mode == 0) {
        *nmx = b == 0 ? c.cx[0] : b == 1 ? c.cx[1] : c.cx[2];
        *nmy = b == 0 ? c.cy[0] : b == 1 ? c.cy[1] : c.cy[2];
        *nref = b == 0 ? c.cr[0] : b == 1 ? c.cr[1] : c.cr[2];
        *ninter = 1;
    } else if (D.mode == 1) { *nmx = D.imx; *nmy = D.imy; *nref = D.iref; *ninter = 1; }
    else if (D.mode == 2)   { *nmx = 0; *nmy = 0; *nref = 0; *ninter = 0; }
    else                    { *nmx = D.qmv[6]; *nmy = D.qmv[7]; *nref = 0; *ninter = 1; }
}

// Chroma of the chosen mode, the reconstruction into the output planes, the
// record for the serialiser and the reported decision.  Returns the luma
// reconstruction (in LDS); the chroma one is in s.crec.
__device__ __forceinline__ const uint8_t *mb_finish(Smem &s, const fdvg_seq &q, int mb, int bx, int by,
                                                    const MbCtx &c, const MbDec &D, fdvg_mbrec *R) {
    const int tid = threadIdx.x, cbx = bx >> 1, cby = by >> 1;
    const double lambda = q.lambda;
    const int qp = q.qp, mode = D.mode;
    int nmx, nmy, nref, ninter;
    mb_motion(c, D, &nmx, &nmy, &nref, &ninter);
    const uint8_t *yrec = mode == 0 ? s.recS : mode == 1 ? s.recI : mode == 2 ? s.recN : s.rec8;
    int ckind = 0;
    if (mode == 0) {
        if (tid < 128) {
            const int pl = tid >> 6, p = tid & 63;
            s.crec[tid] = (uint8_t)chromaq(s.cwin[nref][pl], p & 7, p >> 3, nmx / 2, nmy / 2);
        }
        __syncthreads();
    } else if (mode == 1) {
        if (tid < 128) {
            const int pl = tid >> 6, p = tid & 63;
            s.cpred[tid] = (uint8_t)chromaq(s.cwin[D.iref][pl], p & 7, p >> 3, D.imx / 2, D.imy / 2);
        }
        __syncthreads();
        code_chroma<1>(s, qp, lambda);
        ckind = 1;
    } else if (mode == 2) {
        if (tid < 128) {
            const int pl = tid >> 6, p = tid & 63;
            const uint8_t *tp = &s.ctop[pl][1], *lf = &s.lcol[16 + 8 * pl];
            const Nb nb = nb_setup(tp, lf, s.ctop[pl][0], 8, c.ht, c.hl);
            s.cpred[tid] = (uint8_t)ipred(D.nsub, nb, tp, lf, 8, p & 7, p >> 3);
        }
        __syncthreads();
        code_chroma<0>(s, qp, lambda);
        ckind = 2;
    } else {
        if (tid < 128) {
            const int pl = tid >> 6, p = tid & 63, x = p & 7, y = p >> 3, qd = (y >> 2) * 2 + (x >> 2);
            const int vx = qd == 0 ? D.qmv[0] : qd == 1 ? D.qmv[2] : qd == 2 ? D.qmv[4] : D.qmv[6];
            const int vy = qd == 0 ? D.qmv[1] : qd == 1 ? D.qmv[3] : qd == 2 ? D.qmv[5] : D.qmv[7];
            s.cpred[tid] = (uint8_t)chromaq(s.cwin[0][pl], x, y, vx / 2, vy / 2);
        }
        __syncthreads();
        code_chroma<2>(s, qp, lambda);
        ckind = 3;
    }

    // reconstruction out (the next frame's reference, deblocked later)
    gmemw(q.oy)[(ptrdiff_t)(by + (tid >> 4)) * q.ys + bx + (tid & 15)] = yrec[tid];
    if (tid < 128) {
        const int pl = tid >> 6, p = tid & 63;
        gmemw(pl ? q.ov : q.ou)[(ptrdiff_t)(cby + (p >> 3)) * q.cs + cbx + (p & 7)] = s.crec[tid];
    }
    // decision record for the serialiser
    const int16_t *lsrc = mode == 1 ? s.levI : mode == 2 ? s.levN : s.lev8;
    const uint8_t *fsrc = mode == 1 ? s.flagI : mode == 2 ? s.flagN : s.flag8;
    R->lev[tid] = mode == 0 ? (int16_t)0 : lsrc[tid];
    if (tid < 128) R->clev[tid] = ckind ? s.clev[tid] : (int16_t)0;
    if (tid == 0) {
        const int pmx = c.pmx, pmy = c.pmy;
        R->mode = (uint8_t)mode; R->ref = (uint8_t)(mode == 0 ? nref : mode == 1 ? D.iref : 0);
        R->merge_idx = (uint8_t)D.bmi; R->ncand = (uint8_t)c.ncand;
        R->sub = (uint8_t)D.nsub; R->ckind = (uint8_t)ckind;
        for (int k = 0; k < 4; ++k) R->lflag[k] = mode == 0 ? 0 : fsrc[k];
        R->cflag[0] = s.cflag[0]; R->cflag[1] = s.cflag[1];
        for (int k = 0; k < 4; ++k) { R->mvd[k][0] = R->mvd[k][1] = 0; R->mv[k][0] = R->mv[k][1] = 0; }
        if (mode == 0) { R->mv[0][0] = (int16_t)nmx; R->mv[0][1] = (int16_t)nmy; }
        if (mode == 1) {
            R->mv[0][0] = (int16_t)D.imx; R->mv[0][1] = (int16_t)D.imy;
            R->mvd[0][0] = (int16_t)(D.imx - pmx); R->mvd[0][1] = (int16_t)(D.imy - pmy);
        }
        if (mode == 3) {
            int qpx = pmx, qpy = pmy;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                R->mv[k][0] = (int16_t)D.qmv[2 * k + 0]; R->mv[k][1] = (int16_t)D.qmv[2 * k + 1];
                R->mvd[k][0] = (int16_t)(D.qmv[2 * k + 0] - qpx); R->mvd[k][1] = (int16_t)(D.qmv[2 * k + 1] - qpy);
                qpx = D.qmv[2 * k + 0]; qpy = D.qmv[2 * k + 1];
            }
        }
        if (q.info) {
            gmem_t<fdvg_mbinfo> *I = gmemw(&q.info[mb - q.mb0]);
            I->mode = (uint8_t)mode; I->ref = R->ref; I->merge_idx = (uint8_t)D.bmi; I->sub = (uint8_t)D.nsub;
            for (int k = 0; k < 4; ++k) { I->mv[k][0] = R->mv[k][0]; I->mv[k][1] = R->mv[k][1]; }
        }
    }
    return yrec;
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
        if (row >= nrows) {
            return;
        }
        const int2 rt = rowtab[row];
        const fdvg_seq q = seqs[rt.x];
        const int my = rt.y, mw = q.w >> 4;
        const uint64_t tag = (uint64_t)(q.tag & 0xff) << 56;
        int pmx = 0, pmy = 0;

        for (int mx = 0; mx < mw; ++mx) {
            const int bx = mx * 16, by = my * 16;
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
            mb_stage(s, q, mb, bx, by);
            __syncthreads();

            // ---- neighbours from the row above / the previous MB
            MbCtx c;
            c.ht = my > 0; c.hl = mx > 0;
            c.pmx = pmx; c.pmy = pmy;
            if (tid < 16) s.hbytes[1 + tid] = c.ht ? (uint8_t)hbyte(s, 1, tid) : 0;
            if (tid == 16) s.hbytes[0] = (c.ht && c.hl) ? (uint8_t)hbyte(s, 0, 15) : 128;
            if (tid >= 32 && tid < 50) {
                const int pl = (tid - 32) / 9, k = (tid - 32) % 9;
                s.ctop[pl][k] = k ? (c.ht ? (uint8_t)hbyte(s, 1, 16 + 8 * pl + k - 1) : 0)
                                  : ((c.ht && c.hl) ? (uint8_t)hbyte(s, 0, 16 + 8 * pl + 7) : 128);
            }
            // merge candidates (fdv.h:5711-5732)
            c.cx[0] = pmx; c.cy[0] = pmy; c.cr[0] = 0; c.ncand = 1;
            c.cx[1] = c.cy[1] = c.cr[1] = c.cx[2] = c.cy[2] = c.cr[2] = 0;
            if (c.ht) {
                for (int k = 0; k < 2 && c.ncand < 3; ++k) {
                    const int j = 1 + k;
                    if (mx + k >= mw) continue;
                    const int inter = hbyte(s, j, 37);
                    if (!inter) continue;
                    const int vx = (int16_t)(hbyte(s, j, 32) | (hbyte(s, j, 33) << 8));
                    const int vy = (int16_t)(hbyte(s, j, 34) | (hbyte(s, j, 35) << 8));
                    const int vr = hbyte(s, j, 36);
                    bool dup = false;
#pragma unroll
                    for (int m = 0; m < 3; ++m) dup |= (m < c.ncand && c.cx[m] == vx && c.cy[m] == vy && c.cr[m] == vr);
                    if (!dup) {
                        if (c.ncand == 1) { c.cx[1] = vx; c.cy[1] = vy; c.cr[1] = vr; }
                        else              { c.cx[2] = vx; c.cy[2] = vy; c.cr[2] = vr; }
                        ++c.ncand;
                    }
                }
            }
            __syncthreads();

            const fdvg_mbinfo *gv = q.given ? &q.given[mb - q.mb0] : nullptr;
            const fdvg_metab *mt = (q.me_mode == FDVG_ME_TABLE && q.metab) ? &q.metab[mb - q.mb0] : nullptr;
            MbDec D;
            mb_decide(s, q, c, gv, mt, par, D);
            const uint8_t *yrec = mb_finish(s, q, mb, bx, by, c, D, &recs[mb]);
            int nmx, nmy, nref, ninter;
            mb_motion(c, D, &nmx, &nmy, &nref, &ninter);
            const int mode = D.mode;

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

// ---------------------------------------------------- block-parallel decisions
//
// The parallel and learned modes make the MB loop's own decisions
// (fdv.h:5808-6162) on a motion table, as the exhaustive mode does, but take
// the per-MB coding off the MB chain:
//   k_pcosts   every MB's candidate costs except INTRA, all MBs at once (one
//              workgroup a MB, taken in the replay's order): the SKIP SSDs of
//              likely merge candidates, and INTER16 on each reference and
//              INTER8 coded, with their reconstructions' bottom rows and right
//              columns; published as tagged words;
//   k_pxrep    the reference's decision logic replayed on those costs, one
//              workgroup a MB row: the true vector predictor and merge list,
//              the gates and the strict-< commit, and INTRA costed in the row
//              from the true reconstruction of the MB's left, top and top-left
//              neighbours (the chosen candidates' edges), so that every
//              decision is the MB loop's;
//   k_pencode / k_pintra, or the coding workers of k_pdecide: the decided MBs
//              coded exactly.
// The costs never wait; a replay row waits for its MBs' cost words, for the
// row above (a lower ticket) and nothing else; the coding workers wait only
// for decisions.

__device__ __forceinline__ int mb_count(const fdvg_seq &q) { return (q.w >> 4) * (q.h >> 4); }

#define PC_WORDS 34                      // a MB's costs as tagged words (4 payload bytes each)
#define EW_WORDS 24                      // a MB's candidate edges as tagged words
#define BW_WORDS 8                       // a decided MB's reconstruction edges, tagged: bottom row, right column
static_assert(sizeof(fdvg_pcost) == 4 * PC_WORDS, "cost words");

// The costs of one MB's candidates except INTRA, built in LDS (P), and the
// bottom rows and right columns of its coded candidates' reconstructions in
// sedge: [0..31] INTER16 on reference 0 (bottom row 16, then right column 16),
// [32..63] on reference 1, [64..95] INTER8.  Entries of candidates that are not
// coded (a reference that does not exist, INTER8 below its gate) are never
// read.  q.given must be NULL.
__device__ __forceinline__ void pcost_mb(Smem &s, int32_t (*sred)[4], uint8_t *sedge, const fdvg_seq &q, int i,
                                         fdvg_pcost *P) {
    const int tid = threadIdx.x, wave = tid >> 6;
    const int mw = q.w >> 4, mx = i % mw, my = i / mw, bx = mx * 16, by = my * 16;
    const int mb = q.mb0 + i;
    const gmem_t<fdvg_metab> *T = gmem(&q.metab[i]);
    mb_stage(s, q, mb, bx, by);
    __syncthreads();

    // SKIP SSDs of likely merge candidates: (0,0), and the table vectors of the
    // left, above, above-right (16x16 on both references, last quadrant), above-
    // left and left-left MBs; deduplicated, references that exist only.  The left
    // MB's vector reaches this MB only as the predictor, which is always tried on
    // reference 0, so its reference-1 vector is taken on reference 0.
    typedef int i16v __attribute__((ext_vector_type(16)));
    static_assert(FDVG_NSC <= 16, "candidate vectors");
    int ncs = 0;
    i16v csx = {}, csy = {}, csr = {};
    auto add = [&](int x, int y, int r) {
        if (r >= q.navail || ncs >= FDVG_NSC) return;
        for (int k = 0; k < ncs; ++k) if (csx[k] == x && csy[k] == y && csr[k] == r) return;
        csx[ncs] = x; csy[ncs] = y; csr[ncs] = r; ++ncs;
    };
    add(0, 0, 0);
    const int nbx[5] = {mx - 1, mx, mx + 1, mx - 1, mx - 2}, nby[5] = {my, my - 1, my - 1, my - 1, my};
    for (int k = 0; k < 5; ++k) {
        if (nbx[k] < 0 || nbx[k] >= mw || nby[k] < 0) continue;
        const gmem_t<fdvg_metab> *A = gmem(&q.metab[nby[k] * mw + nbx[k]]);
        add(A->mv16[0][0], A->mv16[0][1], 0);
        if (k < 3) {
            if (q.navail > 1) add(A->mv16[1][0], A->mv16[1][1], k == 0 ? 0 : 1);
            add(A->mv8[3][0], A->mv8[3][1], 0);
        }
    }
    {
        // wave w takes candidates w, w + 4, w + 8; lane L column L & 15, rows 4 (L >> 4) .. + 3
        const int L = tid & 63, x = L & 15, y0 = (L >> 4) * 4;
        for (int c = wave; c < ncs; c += 4) {                 // uniform in the wave
            const int vx = csx[c], vy = csy[c];
            const int e = qerr4<WIN, true>(s.win[csr[c]], WOFF + x + (vx >> 2), WOFF + y0 + (vy >> 2),
                                           ((vy & 3) << 2) | (vx & 3), s.cur + y0 * 16 + x, 16);
            const int t = wave_sum(e);
            if (L == 0) sred[c][0] = t;
        }
    }
    __syncthreads();
    if (tid < FDVG_NSC) {
        const int c = tid;
        P->scx[c] = (int16_t)(c < ncs ? csx[c] : 0);
        P->scy[c] = (int16_t)(c < ncs ? csy[c] : 0);
        P->scr[c] = (uint8_t)(c < ncs ? csr[c] : 0);
        P->sce[c] = c < ncs ? sred[c][0] : 0;
    }
    if (tid == 0) P->nsc = (uint8_t)ncs;
    const int qp = q.qp;
    const double lambda = q.lambda;
    // INTER16 on each reference
    for (int r = 0; r < q.navail; ++r) {
        s.pred[tid] = (uint8_t)lumaq(s.win[r], tid & 15, tid >> 4, T->mv16[r][0], T->mv16[r][1]);
        __syncthreads();
        code_luma(s, s.pred, s.recT, s.levT, s.flagT, qp, lambda);
        const int d = s.ws.e[0] + s.ws.e[1] + s.ws.e[2] + s.ws.e[3];
        const int b = 6 + s.ws.b[0] + s.ws.b[1] + s.ws.b[2] + s.ws.b[3];
        if (tid == 0) { P->d16[r] = d; P->b16[r] = b; }
        if (tid < 32) sedge[32 * r + tid] = tid < 16 ? s.recT[15 * 16 + tid] : s.recT[(tid - 16) * 16 + 15];
        __syncthreads();
    }
    // INTER8 is evaluated only when the 16x16 SAD of the chosen reference (either
    // table SAD) exceeds the quadrants' sum by more than 10% (fdv.h:6014); where
    // neither does, its cost is never read.  Also not where the proposer expects
    // SKIP.
    {
        const int ssad = T->sad8[0] + T->sad8[1] + T->sad8[2] + T->sad8[3];
        const bool g0 = T->sad16[0] > 0 && ssad * 10 >= T->sad16[0] * 9;
        const bool g1 = q.navail < 2 || (T->sad16[1] > 0 && ssad * 10 >= T->sad16[1] * 9);
        if ((T->flags & 1) || (g0 && g1)) {
            if (tid == 0) P->d8 = P->b8 = 0;
            return;
        }
    }
    // INTER8 from the quadrant vectors
    {
        const int x = tid & 15, y = tid >> 4, qd = (y >> 3) * 2 + (x >> 3);
        s.pred[tid] = (uint8_t)lumaq(s.win[0], x, y, T->mv8[qd][0], T->mv8[qd][1]);
    }
    __syncthreads();
    code_luma(s, s.pred, s.rec8, s.lev8, s.flag8, qp, lambda);
    if (tid == 0) {
        P->d8 = s.ws.e[0] + s.ws.e[1] + s.ws.e[2] + s.ws.e[3];
        P->b8 = 3 + s.ws.b[0] + s.ws.b[1] + s.ws.b[2] + s.ws.b[3];
    }
    if (tid < 32) sedge[64 + tid] = tid < 16 ? s.rec8[15 * 16 + tid] : s.rec8[(tid - 16) * 16 + 15];
}

// LDS of one MB's costs (k_pcosts).
struct CostLds {
    Smem s;
    int32_t sred[FDVG_NSC][4];
    uint8_t sedge[96];
    union { fdvg_pcost c; uint32_t w[PC_WORDS]; } sp;
};

// The costs of every MB, one workgroup a MB, published as tagged words: the
// costs (PC_WORDS a MB) and the candidate edges (EW_WORDS a MB, 4 samples a
// word), read by the replays, which may run at the same time.  XCD k (the
// workgroups with blockIdx.x % 8 == k) takes, by its own ticket (ctl[8 + k]),
// the MBs of its rows in the replays' order (xorder[xoff[k] .. xoff[k + 1]]):
// the windows that neighbouring MBs share are then fetched into one L2, and
// the costs still arrive in about the order the replays use them.
extern "C" __global__ void __launch_bounds__(256, 4)
k_pcosts(const fdvg_seq *__restrict__ seqs, int nseq, const unsigned *__restrict__ xorder,
         const int *__restrict__ xoff, unsigned *ctl, uint64_t *__restrict__ pcw, uint64_t *__restrict__ edgew) {
    __shared__ CostLds C;
    __shared__ int job_s;
    const int tid = threadIdx.x, xk = blockIdx.x & 7;
    if (tid == 0) job_s = (int)atomicAdd(ctl + 8 + xk, 1u);
    __syncthreads();
    const int job = __builtin_amdgcn_readfirstlane(job_s);
    const int lo = xoff[xk], n = xoff[xk + 1] - lo;
    if (job >= n) return;
    const int mb = __builtin_amdgcn_readfirstlane((int)xorder[lo + job]);
    int sq = 0;
    for (int k = 1; k < nseq; ++k) if (mb >= seqs[k].mb0) sq = k;
    fdvg_seq q = seqs[sq];
    q.given = nullptr;
    pcost_mb(C.s, C.sred, C.sedge, q, mb - q.mb0, &C.sp.c);
    __syncthreads();
    const uint64_t tag = (uint64_t)q.step << 32;
    if (tid < EW_WORDS) {
        const uint32_t v = (uint32_t)C.sedge[4 * tid] | (uint32_t)C.sedge[4 * tid + 1] << 8 |
                           (uint32_t)C.sedge[4 * tid + 2] << 16 | (uint32_t)C.sedge[4 * tid + 3] << 24;
        __hip_atomic_store(edgew + (size_t)mb * EW_WORDS + tid, tag | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (tid >= 64 && tid < 64 + PC_WORDS) {
        __hip_atomic_store(pcw + (size_t)mb * PC_WORDS + tid - 64, tag | C.sp.w[tid - 64], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Motion record handoff of the decision wavefront: tag | inter | ref | x | y.
__device__ __forceinline__ uint64_t mrec_pack(uint64_t tag, int inter, int ref, int x, int y) {
    return tag | ((uint64_t)(inter & 1) << 33) | ((uint64_t)(ref & 1) << 32) |
           ((uint64_t)(uint16_t)x << 16) | (uint64_t)(uint16_t)y;
}

// SSD of the 16x16 block `curb` against its prediction from a tile (row stride
// S, the block's (0, 0) at tile (2, 2)) at quarter-pel phase ph, by one
// wavefront: lane L takes column L & 15, rows 4 (L >> 4) .. + 3 (qerr4).
template <int S>
__device__ __forceinline__ int ssd_tile_q(const uint8_t *tile, const uint8_t *curb, int ph, int L) {
    const int x = L & 15, y0 = (L >> 4) * 4;
    return qerr4<S, true, true>(tile, x + 2, y0 + 2, ph, curb + y0 * 16 + x, 16);
}

// The replay wave's staging of a row, PR_CHUNK MBs at a time: costs, candidate
// edges, table entries and source MBs.
#define PR_CHUNK 8
#define PR_TS 24                         // tile stride: 21 samples a row
struct __align__(16) PrWave {
    fdvg_pcost pc[PR_CHUNK];
    fdvg_metab mt[PR_CHUNK];
    uint32_t ew[PR_CHUNK][EW_WORDS];     // candidate edges, 4 samples a word
    uint8_t cur[PR_CHUNK][256];          // the source MBs, for SSDs outside the precomputed set
    uint8_t tile[21 * PR_TS];
};

// A chunk of a row in flight: every lane's share of the costs, edges, tables
// and source MBs, fixed trip counts so the loads issue back to back.
#define PR_PW ((PR_CHUNK * PC_WORDS + 63) / 64)
#define PR_EW ((PR_CHUNK * EW_WORDS + 63) / 64)
#define PR_MW ((PR_CHUNK * (int)sizeof(fdvg_metab) / 4 + 63) / 64)
#define PR_CW (PR_CHUNK * 64 / 64)
struct PrPre {
    uint64_t p[PR_PW], e[PR_EW];         // tagged cost and edge words
    uint32_t m[PR_MW], c[PR_CW];
    int n;
};
// Issues the chunk's loads; the cost and edge words may still be in flight
// (k_pcosts running concurrently): pr_store waits for their tags.
__device__ __forceinline__ void pr_load(PrPre &R, const fdvg_seq &q, const uint64_t *pcw, const uint64_t *edgew,
                                        int my, int mx0, int L) {
    const int mw = q.w >> 4, i0 = my * mw + mx0;
    const int n = mw - mx0 < PR_CHUNK ? mw - mx0 : PR_CHUNK;
    const uint64_t *gp = pcw + (size_t)(q.mb0 + i0) * PC_WORDS, *ge = edgew + (size_t)(q.mb0 + i0) * EW_WORDS;
    const gmem_t<uint32_t> *gm = gmem((const uint32_t *)(q.metab + i0));
    const gmem_t<uint8_t> *gc = gmem(q.cy);
    const int np = n * PC_WORDS, ne = n * EW_WORDS, nm = n * (int)(sizeof(fdvg_metab) / 4);
    const uint64_t none = (uint64_t)q.step << 32;
#pragma unroll
    for (int j = 0; j < PR_PW; ++j) {
        const int k = L + 64 * j;
        R.p[j] = k < np ? __hip_atomic_load(gp + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : none;
    }
#pragma unroll
    for (int j = 0; j < PR_EW; ++j) {
        const int k = L + 64 * j;
        R.e[j] = k < ne ? __hip_atomic_load(ge + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : none;
    }
#pragma unroll
    for (int j = 0; j < PR_MW; ++j) { const int k = L + 64 * j; R.m[j] = k < nm ? gm[k] : 0u; }
#pragma unroll
    for (int j = 0; j < PR_CW; ++j) {
        const int k = L + 64 * j, m = k >> 6, rr = (k >> 2) & 15, cc = (k & 3) * 4;
        R.c[j] = m < n ? *(const gmem_t<uint32_t> *)(gc + (size_t)(my * 16 + rr) * q.w + (mx0 + m) * 16 + cc) : 0u;
    }
    R.n = n;
}
// Waits until every cost and edge word of the chunk carries this frame step's
// id (reloading the stale ones; bounded, error flag on timeout), then stores
// the chunk into the wave's LDS block.
__device__ __forceinline__ void pr_store(PrWave &W, PrPre &R, const fdvg_seq &q, const uint64_t *pcw,
                                         const uint64_t *edgew, unsigned *ctl, int my, int mx0, int L) {
    const int mw = q.w >> 4, i0 = my * mw + mx0;
    const uint64_t *gp = pcw + (size_t)(q.mb0 + i0) * PC_WORDS, *ge = edgew + (size_t)(q.mb0 + i0) * EW_WORDS;
    const int np = R.n * PC_WORDS, ne = R.n * EW_WORDS, nm = R.n * (int)(sizeof(fdvg_metab) / 4);
    bool ok = true;
#pragma unroll
    for (int j = 0; j < PR_PW; ++j) ok = ok && (uint32_t)(R.p[j] >> 32) == q.step;
#pragma unroll
    for (int j = 0; j < PR_EW; ++j) ok = ok && (uint32_t)(R.e[j] >> 32) == q.step;
    for (unsigned spins = 0; !__all(ok); ++spins) {
        if (spins > (1u << 22)) { if (L == 0) atomicOr(ctl + 1, 1u); break; }
        __builtin_amdgcn_s_sleep(1);
        ok = true;
#pragma unroll
        for (int j = 0; j < PR_PW; ++j) {
            const int k = L + 64 * j;
            if (k < np && (uint32_t)(R.p[j] >> 32) != q.step)
                R.p[j] = __hip_atomic_load(gp + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = ok && (uint32_t)(R.p[j] >> 32) == q.step;
        }
#pragma unroll
        for (int j = 0; j < PR_EW; ++j) {
            const int k = L + 64 * j;
            if (k < ne && (uint32_t)(R.e[j] >> 32) != q.step)
                R.e[j] = __hip_atomic_load(ge + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = ok && (uint32_t)(R.e[j] >> 32) == q.step;
        }
    }
    uint32_t *lp = (uint32_t *)W.pc, *le = &W.ew[0][0], *lm = (uint32_t *)W.mt, *lc = (uint32_t *)W.cur;
#pragma unroll
    for (int j = 0; j < PR_PW; ++j) { const int k = L + 64 * j; if (k < np) lp[k] = (uint32_t)R.p[j]; }
#pragma unroll
    for (int j = 0; j < PR_EW; ++j) { const int k = L + 64 * j; if (k < ne) le[k] = (uint32_t)R.e[j]; }
#pragma unroll
    for (int j = 0; j < PR_MW; ++j) { const int k = L + 64 * j; if (k < nm) lm[k] = R.m[j]; }
#pragma unroll
    for (int j = 0; j < PR_CW; ++j) lc[L + 64 * j] = R.c[j];
}

// ---- the exact replay (k_pxrep)
//
// One workgroup of XR_WAVES waves replays one MB row.  Wave 0 runs the
// reference's decision logic (fdv.h:5847-6040) on the stored costs, MB by MB;
// waves 1-4 are the INTRA trial: for MB x they screen the four sub-modes by
// SATD from the true reconstruction around it (the row above's bottom rows,
// published by the row above's workgroup; the left MB's right column, just
// decided by wave 0) and code the chosen prediction, one 8x8 region a wave,
// while wave 0 weighs the other candidates.  When wave 0 does not consult
// INTRA for an MB it tells the trial to stop (`cancel`).  After each decision
// wave 0 forms the chosen candidate's bottom row and right column (INTER16 and
// INTER8: from k_pcosts; SKIP: the prediction itself, from a tile of the
// reference; INTRA: the trial's reconstruction): the right column is the next
// MB's left neighbour, the bottom row goes to the row below as tagged words
// (BW_WORDS a MB: 4 of the bottom row, then 4 of the right column, which the
// coding workers read for an intra MB's luma prediction).  The trial's inputs are double-buffered by MB parity; wave 0 refills
// a buffer only once the trial has finished the MB that used it.
#define XR_WAVES 5
struct __align__(16) XrLds {
    PrWave pw;
    uint32_t itop[2][4];                  // per MB parity: the row above, 4 samples a word
    uint32_t ileft[2][4];                 // per MB parity: the left MB's right column
    int icorner[2];                       // per MB parity: the above-left sample
    uint8_t icur[4][256];                 // each trial wave's copy of the source MB
    uint8_t ipred[256], irec[256];
    int16_t ilev[256];
    uint8_t iflag[4];
    RegionWs ws;
    Tab8 tab;                             // the 8x8 basis for the trial's region coding
    int sat[2][4];                        // per MB parity: the SATD of each sub-mode
    int sdone[4];                         // per trial wave: MBs whose SATD it has published
    int go;                               // MB x's trial inputs are ready once go > x
    int cancel[2];                        // per MB parity: x + 1 when MB x's INTRA cost is not wanted
    int done[4];                          // per trial wave: MBs finished
};

__device__ __forceinline__ int lds_acq(int *p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP); }
// A bounded wait gives up after `limit` spins, or sooner once another wait
// has given up (the error flag ctl[1], read every 256 spins): a failed step
// then ends quickly instead of every later wait running to its own limit.
__device__ __forceinline__ bool xr_give_up(unsigned spins, unsigned limit, unsigned *ctl) {
    if (spins > limit) { atomicOr(ctl + 1, 1u); return true; }
    return (spins & 255u) == 255u && __hip_atomic_load(ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}
__device__ __forceinline__ void lds_rel(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP); }

// Wave 0 waits until the trial waves have finished MB x - 1 (done >= x).
__device__ __forceinline__ void xr_wait_done(XrLds &X, int x, unsigned *ctl, int L) {
    int v = L < 4 ? lds_acq(&X.done[L]) : x;
    for (unsigned spins = 0; !__all(v >= x); ++spins) {
        if (xr_give_up(spins, 1u << 24, ctl)) break;
        __builtin_amdgcn_s_sleep(1);
        if (L < 4) v = lds_acq(&X.done[L]);
    }
}

#define PD_WORDS 10
__device__ __forceinline__ uint32_t pk16(int lo, int hi) { return (uint32_t)(uint16_t)lo | (uint32_t)(uint16_t)hi << 16; }

// Wave 0 of k_pxrep: the decisions of row my.  Publishes each decision as
// given[] / pctx[] (for the separate coding launches), as ten tagged words in
// decw (for the coding workers of k_pdecide), and as a tagged motion record in
// mrec (for the row below), followed by the bottom row in botw.
__device__ __forceinline__ void xr_decide(XrLds &X, const fdvg_seq &q, const uint64_t *__restrict__ pcw,
                                          const uint64_t *__restrict__ edgew, unsigned *ctl, int my,
                                          uint64_t *__restrict__ mrec, uint64_t *__restrict__ botw,
                                          uint64_t *__restrict__ decw, fdvg_mbinfo *__restrict__ given,
                                          fdvg_pctx *__restrict__ pctx, int L) {
    PrWave &W = X.pw;
    const int mw = q.w >> 4;
    const double lambda = q.lambda;
    const uint64_t tagv = (uint64_t)(q.step & 0x3fffffffu) << 34;
    const unsigned step = q.step;
    const uint64_t *arec = my > 0 ? mrec + q.mb0 + (size_t)(my - 1) * mw : nullptr;
    const uint64_t *abot = my > 0 ? botw + (size_t)(q.mb0 + (my - 1) * mw) * BW_WORDS : nullptr;
    // The row above, lane k < 4: bottom word k of an above MB; lanes 4, 5:
    // motion records.  Waits (bounded) for the words of this step.
    auto above = [&](int ib, int ir0, int ir1) -> uint64_t {
        const uint64_t *p = L < 4 ? abot + (size_t)ib * BW_WORDS + L : L == 4 ? arec + ir0 : arec + ir1;
        const bool act = L < 6 && (L < 4 ? ib >= 0 : true);
        uint64_t v = act ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        auto fresh = [&](uint64_t w) { return !act || (L < 4 ? (uint32_t)(w >> 32) == step : (w >> 34) == (tagv >> 34)); };
        for (unsigned spins = 0; !__all(fresh(v)); ++spins) {
            if (xr_give_up(spins, 1u << 22, ctl)) break;
            __builtin_amdgcn_s_sleep(1);
            if (!fresh(v)) v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return v;
    };
    // MB 0's trial inputs
    uint64_t ab = 0;
    if (my > 0) ab = above(0, 0, mw > 1 ? 1 : 0);
    if (L < 4) X.itop[0][L] = (uint32_t)ab;
    if (L == 0) X.icorner[0] = 128;
    wave_sync();
    if (L == 0) lds_rel(&X.go, 1);

    int pmx = 0, pmy = 0;
    PrPre pre;
    for (int mx = 0; mx < mw; ++mx) {
        const int i = my * mw + mx, mb = q.mb0 + i, bx = mx * 16, by = my * 16;
        if (mx % PR_CHUNK == 0) {
            if (mx == 0) pr_load(pre, q, pcw, edgew, my, 0, L);
            pr_store(W, pre, q, pcw, edgew, ctl, my, mx, L);
            wave_sync();
            if (mx + PR_CHUNK < mw) pr_load(pre, q, pcw, edgew, my, mx + PR_CHUNK, L);
        }
        const fdvg_pcost &P = W.pc[mx % PR_CHUNK];
        const fdvg_metab &T = W.mt[mx % PR_CHUNK];
        const int nsc = P.nsc;
        const int sx_l = L < FDVG_NSC ? P.scx[L] : 0, sy_l = L < FDVG_NSC ? P.scy[L] : 0;
        const int sr_l = L < FDVG_NSC ? P.scr[L] : 0, se_l = L < FDVG_NSC ? P.sce[L] : 0;
        const int d16_0 = P.d16[0], d16_1 = P.d16[1], b16_0 = P.b16[0], b16_1 = P.b16[1];
        const int pd8 = P.d8, pb8 = P.b8;
        const int m0x = T.mv16[0][0], m0y = T.mv16[0][1], m1x = T.mv16[1][0], m1y = T.mv16[1][1];
        const int s16_0 = T.sad16[0], s16_1 = T.sad16[1], tfl = T.flags;
        const bool pruned = tfl & 1;          // the proposer expects SKIP: no INTRA, no INTER8
        if (pruned && L == 0) lds_rel(&X.cancel[mx & 1], mx + 1);
        int q8x[4], q8y[4];
#pragma unroll
        for (int qd = 0; qd < 4; ++qd) { q8x[qd] = T.mv8[qd][0]; q8y[qd] = T.mv8[qd][1]; }
        const int ssad = T.sad8[0] + T.sad8[1] + T.sad8[2] + T.sad8[3];
        // merge candidate 0 is the predictor on reference 0, known already:
        // outside the precomputed set, its 21x21 tile loads are issued now so
        // that they complete during the wait below
        const bool hit0 = __ballot(L < nsc && sx_l == pmx && sy_l == pmy && sr_l == 0) != 0;
        uint8_t tv0[7];
        if (!hit0) {
            const gmem_t<uint8_t> *rp = gmem(q.ry[0]) + (ptrdiff_t)(by + (pmy >> 2) - 2) * q.ys + bx + (pmx >> 2) - 2;
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int k = L + 64 * j;
                tv0[j] = k < 21 * 21 ? rp[(ptrdiff_t)(k / 21) * q.ys + k % 21] : 0;
            }
        }
        // the row above: the motion records of the above and above-right MBs
        // (or the last), and the bottom row of the above-right MB (the next
        // MB's trial)
        uint64_t va = 0, vr = 0, nb = 0;
        if (my > 0) {
            nb = above(mx + 1 < mw ? mx + 1 : -1, mx, mx + 1 < mw ? mx + 1 : mw - 1);
            va = __shfl(nb, 4);
            vr = mx + 1 < mw ? __shfl(nb, 5) : 0;
        }
        // merge list: predictor, then the inter MBs above and above-right (fdv.h:5711-5732)
        int cx0 = pmx, cy0 = pmy, cx1 = 0, cy1 = 0, cr1 = 0, cx2 = 0, cy2 = 0, cr2 = 0, ncand = 1;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint64_t w = k ? vr : va;
            if (!((w >> 33) & 1) || ncand >= 3) continue;
            const int x = (int16_t)((w >> 16) & 0xffff), y = (int16_t)(w & 0xffff), r = (int)((w >> 32) & 1);
            const bool dup = (cx0 == x && cy0 == y && r == 0) || (ncand > 1 && cx1 == x && cy1 == y && cr1 == r);
            if (dup) continue;
            if (ncand == 1) { cx1 = x; cy1 = y; cr1 = r; } else { cx2 = x; cy2 = y; cr2 = r; }
            ++ncand;
        }
        // SKIP SSDs: from the precomputed set, else from a 21x21 tile of the reference
        int ssd[3] = {0, 0, 0};
        int tile_c = -1;                      // the candidate whose tile W.tile holds
        const uint8_t *curb = W.cur[mx % PR_CHUNK];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (c >= ncand) continue;
            const int x = c == 0 ? cx0 : c == 1 ? cx1 : cx2, y = c == 0 ? cy0 : c == 1 ? cy1 : cy2;
            const int r = c == 0 ? 0 : c == 1 ? cr1 : cr2;
            const bool hit = L < nsc && sx_l == x && sy_l == y && sr_l == r;
            const uint64_t hm = __ballot(hit);
            if (hm) {
                ssd[c] = __builtin_amdgcn_readlane(se_l, __builtin_ctzll(hm));
            } else {
                uint8_t tv[7];
                if (c == 0) {
#pragma unroll
                    for (int j = 0; j < 7; ++j) tv[j] = tv0[j];
                } else {
                    const gmem_t<uint8_t> *rp = gmem(r ? q.ry[1] : q.ry[0]) +
                                        (ptrdiff_t)(by + (y >> 2) - 2) * q.ys + bx + (x >> 2) - 2;
#pragma unroll
                    for (int j = 0; j < 7; ++j) {
                        const int k = L + 64 * j;
                        tv[j] = k < 21 * 21 ? rp[(ptrdiff_t)(k / 21) * q.ys + k % 21] : 0;
                    }
                }
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const int k = L + 64 * j;
                    if (k < 21 * 21) W.tile[(k / 21) * PR_TS + k % 21] = tv[j];
                }
                wave_sync();
                ssd[c] = wave_sum(ssd_tile_q<PR_TS>(W.tile, curb, ((y & 3) << 2) | (x & 3), L));
                tile_c = c;
                wave_sync();
            }
        }
        int Ds = ssd[0], bmi = 0;
        if (ncand > 1 && ssd[1] < Ds) { Ds = ssd[1]; bmi = 1; }
        if (ncand > 2 && ssd[2] < Ds) { Ds = ssd[2]; bmi = 2; }
        const int sx = bmi == 0 ? cx0 : bmi == 1 ? cx1 : cx2, sy = bmi == 0 ? cy0 : bmi == 1 ? cy1 : cy2;
        const int sref = bmi == 0 ? 0 : bmi == 1 ? cr1 : cr2;
        // the SKIP winner's tile, for its edges should SKIP be chosen
        const bool need_tile = tile_c != bmi;
        uint8_t tvs[7];
        if (need_tile) {
            const gmem_t<uint8_t> *rp = gmem(sref ? q.ry[1] : q.ry[0]) +
                                        (ptrdiff_t)(by + (sy >> 2) - 2) * q.ys + bx + (sx >> 2) - 2;
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int k = L + 64 * j;
                tvs[j] = k < 21 * 21 ? rp[(ptrdiff_t)(k / 21) * q.ys + k % 21] : 0;
            }
        }
        // the reference's gates and commit on the stored costs
        const double Js = (double)Ds + lambda * (double)(3 + (ncand > 1 ? 3 : 0));
        double Jbest = Js;
        const bool skip_wins = Js <= lambda * (double)FL_INTRA;
        double Ji = 1e30;
        int iref = 0, sad16 = 0;
        if (!skip_wins) {
            const int bb = b16_0 + egbits(zmap(m0x - pmx)) + egbits(zmap(m0y - pmy));
            const double J = (double)d16_0 + lambda * (double)bb;
            if (J < Jbest) Jbest = J;
            Ji = J; sad16 = s16_0;
            if (q.navail > 1 && !(Jbest <= lambda * (double)FL_INTER16)) {
                const int bb1 = b16_1 + egbits(zmap(m1x - pmx)) + egbits(zmap(m1y - pmy));
                const double J1 = (double)d16_1 + lambda * (double)bb1;
                if (J1 < Jbest) Jbest = J1;
                if (J1 < Ji) { Ji = J1; iref = 1; sad16 = s16_1; }
            }
        }
        // the bottom row (lanes 0-15) and right column (lanes 16-31) of each
        // candidate's reconstruction but INTRA's, formed while the trial runs:
        // SKIP's is the prediction itself, from the winner's tile
        if (need_tile) {
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int k = L + 64 * j;
                if (k < 21 * 21) W.tile[(k / 21) * PR_TS + k % 21] = tvs[j];
            }
            wave_sync();
        }
        int e_s = 0, e_i = 0, e_8 = 0;
        if (L < 32) {
            const int ex = L < 16 ? L : 15, ey = L < 16 ? 15 : L - 16;
            const uint32_t *ew = W.ew[mx % PR_CHUNK];
            e_s = lumaq_at<PR_TS>(W.tile, 2 + ex, 2 + ey, ((sy & 3) << 2) | (sx & 3));
            e_i = (int)((ew[(32 * iref + L) >> 2] >> (8 * (L & 3))) & 0xffu);
            e_8 = (int)((ew[(64 + L) >> 2] >> (8 * (L & 3))) & 0xffu);
        }
        double Jn = 1e30;
        if (!skip_wins && !pruned && Jbest > lambda * (double)FL_INTRA) {
            // the trial of this MB: its four region costs
            xr_wait_done(X, mx + 1, ctl, L);
            const int dn = X.ws.e[0] + X.ws.e[1] + X.ws.e[2] + X.ws.e[3];
            const int bn = 6 + X.ws.b[0] + X.ws.b[1] + X.ws.b[2] + X.ws.b[3];
            Jn = (double)dn + lambda * (double)bn;
            if (Jn < Jbest) Jbest = Jn;
        } else if (L == 0) {
            lds_rel(&X.cancel[mx & 1], mx + 1);
        }
        double J8 = 1e30;
        if (!skip_wins && !pruned && Jbest > lambda * (double)FL_INTER8) {
            if (!(sad16 > 0 && ssad * 10 >= sad16 * 9)) {
                int b8 = pb8, qpx = pmx, qpy = pmy;
#pragma unroll
                for (int qd = 0; qd < 4; ++qd) {
                    b8 += egbits(zmap(q8x[qd] - qpx)) + egbits(zmap(q8y[qd] - qpy));
                    qpx = q8x[qd]; qpy = q8y[qd];
                }
                J8 = (double)pd8 + lambda * (double)b8;
            }
        }
        int mode = 0;
        double Jm = Js;
        if (Ji < Jm) { Jm = Ji; mode = 1; }
        if (Jn < Jm) { Jm = Jn; mode = 2; }
        if (J8 < Jm) { Jm = J8; mode = 3; }
        int nx = 0, ny = 0, nr = 0, ni = 1;
        if (mode == 0) { nx = sx; ny = sy; nr = sref; }
        else if (mode == 1) { nx = iref ? m1x : m0x; ny = iref ? m1y : m0y; nr = iref; }
        else if (mode == 2) ni = 0;
        else { nx = q8x[3]; ny = q8y[3]; }
        // the chosen reconstruction's edges
        int e = mode == 0 ? e_s : mode == 1 ? e_i : e_8;
        if (mode == 2 && L < 32) e = X.irec[(L < 16 ? 15 : L - 16) * 16 + (L < 16 ? L : 15)];
        // the next MB's trial inputs, into the buffer of MB mx - 1's parity once
        // the trial has finished that MB: this is the chain, everything else
        // is published after it
        if (mx + 1 < mw) {
            const int p = (mx + 1) & 1;
            xr_wait_done(X, mx, ctl, L);
            if (L >= 16 && L < 32) ((uint8_t *)X.ileft[p])[L - 16] = (uint8_t)e;
            if (L < 4) X.itop[p][L] = (uint32_t)nb;
            if (L == 0) X.icorner[p] = (int)(X.itop[mx & 1][3] >> 24);
            wave_sync();
            if (L == 0) lds_rel(&X.go, mx + 2);
        }
        // the decision: the coding workers' copy (ten tagged words, lane k:
        // payload word k), the bottom row and motion record for the row below,
        // given / pctx
        {
            const uint32_t mvw0 = mode == 3 ? pk16(q8x[0], q8y[0]) : mode <= 1 ? pk16(nx, ny) : 0u;
            uint32_t d = 0;
            switch (L) {
            case 0: d = (uint32_t)mode | (uint32_t)(mode <= 1 ? nr : 0) << 8 | (uint32_t)(mode == 0 ? bmi : 0) << 16 |
                        (uint32_t)(mode == 2 ? FDVG_SUB_SCREEN : 0) << 24; break;
            case 1: d = mvw0; break;
            case 2: d = mode == 3 ? pk16(q8x[1], q8y[1]) : 0u; break;
            case 3: d = mode == 3 ? pk16(q8x[2], q8y[2]) : 0u; break;
            case 4: d = mode == 3 ? pk16(q8x[3], q8y[3]) : 0u; break;
            case 5: d = pk16(pmx, pmy); break;
            case 6: d = pk16(cx0, cx1); break;
            case 7: d = pk16(cx2, cy0); break;
            case 8: d = pk16(cy1, cy2); break;
            case 9: d = (uint32_t)cr1 << 8 | (uint32_t)cr2 << 16 | (uint32_t)ncand << 24; break;
            default: break;
            }
            if (L < PD_WORDS) __hip_atomic_store(decw + (size_t)mb * PD_WORDS + L, (uint64_t)step << 32 | d,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // lanes 0-3: the bottom row (e of lanes 0-15), 4-7: the right column (16-31)
            const int eb = (L & 4) * 4 + 4 * (L & 3);
            const int b0 = __shfl(e, eb), b1 = __shfl(e, eb + 1), b2 = __shfl(e, eb + 2), b3 = __shfl(e, eb + 3);
            if (L < BW_WORDS)
                __hip_atomic_store(botw + (size_t)mb * BW_WORDS + L,
                                   (uint64_t)step << 32 | (uint32_t)(b0 | b1 << 8 | b2 << 16 | b3 << 24),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (L == 0) __hip_atomic_store(mrec + mb, mrec_pack(tagv, ni, nr, nx, ny), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        }
        if (L == 0) {
            fdvg_mbinfo G;
            G.mode = (uint8_t)mode; G.ref = (uint8_t)(mode == 0 || mode == 1 ? nr : 0);
            G.merge_idx = (uint8_t)(mode == 0 ? bmi : 0); G.sub = (uint8_t)(mode == 2 ? FDVG_SUB_SCREEN : 0);
            for (int k = 0; k < 4; ++k) { G.mv[k][0] = 0; G.mv[k][1] = 0; }
            if (mode == 0 || mode == 1) { G.mv[0][0] = (int16_t)nx; G.mv[0][1] = (int16_t)ny; }
            if (mode == 3) for (int k = 0; k < 4; ++k) { G.mv[k][0] = (int16_t)q8x[k]; G.mv[k][1] = (int16_t)q8y[k]; }
            given[mb] = G;
            fdvg_pctx Xc;
            Xc.pmx = (int16_t)pmx; Xc.pmy = (int16_t)pmy; Xc.ncand = (uint8_t)ncand;
            Xc.cx[0] = (int16_t)cx0; Xc.cy[0] = (int16_t)cy0; Xc.cr[0] = 0;
            Xc.cx[1] = (int16_t)cx1; Xc.cy[1] = (int16_t)cy1; Xc.cr[1] = (uint8_t)cr1;
            Xc.cx[2] = (int16_t)cx2; Xc.cy[2] = (int16_t)cy2; Xc.cr[2] = (uint8_t)cr2;
            pctx[mb] = Xc;
        }
        if (mode != 2) { pmx = nx; pmy = ny; }
    }
}

// The neighbour parameters of the 16x16 intra predictions (nb_setup for n = 16)
// with 16-lane DPP sums: lane i of each 16-lane row takes top[i] and left[i];
// PLANE's H = sum (i - 7) top[i] - 8 corner, V likewise.
__device__ __forceinline__ Nb xr_nb(const uint8_t *top, const uint8_t *left, int tl, int ht, int hl, int L) {
    const int i = L & 15;
    const int ti = ht ? top[i] : 0, li = hl ? left[i] : 0;
    const int st = row_sum16(ti), sl = row_sum16(li);
    Nb q;
    q.ht = ht; q.hl = hl;
    q.dc = ht && hl ? (st + sl + 16) / 32 : ht ? (st + 8) / 16 : hl ? (sl + 8) / 16 : 128;
    q.a = q.b = q.c = 0;
    if (ht && hl) {
        const int H = row_sum16((i - 7) * ti) - 8 * tl, V = row_sum16((i - 7) * li) - 8 * tl;
        q.a = 16 * (left[15] + top[15]);
        q.b = (5 * H + 32) >> 6;
        q.c = (5 * V + 32) >> 6;
    }
    return q;
}

// The SATD of sub-mode m over the MB (mb_satd, fdv.h:4411; the 4x4 Hadamard of
// fdv.h:3803-3822), one wavefront: lane L takes row L & 3 of 4x4 block L >> 2,
// transforms it in registers, and the columns across the quad's four lanes by
// DPP (lane j: the sum with lane j ^ s, or the other minus its own value when
// bit s of j is set), which leaves each quad holding the block's 16 Hadamard
// coefficients in some order.  Returns the total (uniform).
__device__ __forceinline__ int xr_satd(const uint8_t *cur, const uint8_t *top, const uint8_t *left, const Nb &q,
                                       int m, int L) {
    const int b = L >> 2, j = L & 3, ox = (b & 3) * 4, y = (b >> 2) * 4 + j;
    const uint32_t cw = *(const uint32_t *)(cur + y * 16 + ox), tw = *(const uint32_t *)(top + ox);
    const int lft = left[y];
    int d[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int x = ox + k;
        int p = q.dc;
        if (m == 3 && q.ht && q.hl) p = clip8((q.a + q.c * (y - 7) + 16 + q.b * (x - 7)) >> 5);
        else if (m == 1 && q.ht) p = (int)((tw >> (8 * k)) & 255u);
        else if (m == 2 && q.hl) p = lft;
        d[k] = (int)((cw >> (8 * k)) & 255u) - p;
    }
    const int p0 = d[0] + d[1], m0 = d[0] - d[1], p1 = d[2] + d[3], m1 = d[2] - d[3];
    int t[4] = {p0 + p1, m0 + m1, p0 - p1, m0 - m1};
    int sa = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int v = t[k];
        int o = dpp_mov<DPP_XOR1>(v);
        v = (j & 1) ? o - v : v + o;
        o = dpp_mov<DPP_XOR2>(v);
        v = (j & 2) ? o - v : v + o;
        sa += abs(v);
    }
    return wave_sum(sa);
}

// Waves 1-4 of k_pxrep: the INTRA trial of every MB of row my, region r of
// each (fdv.h:5950-5988: the SATD screen over DC / V / H / PLANE, then one
// coding of the best), unless wave 0 cancels it.
__device__ __forceinline__ void xr_trial(XrLds &X, const fdvg_seq &q, unsigned *ctl, int my, int r, int L) {
    const int mw = q.w >> 4;
    uint8_t *cur = X.icur[r];
    const gmem_t<uint32_t> *src = (const gmem_t<uint32_t> *)gmem(q.cy);
    const size_t srow = (size_t)(my * 16 + (L >> 2)) * (q.w >> 2) + (L & 3);
    uint32_t cnext = src[srow];
    const int ox = (r & 1) * 8, oy = (r >> 1) * 8, o = oy * 16 + ox;
    for (int mx = 0; mx < mw; ++mx) {
        const uint32_t cv = cnext;
        if (mx + 1 < mw) cnext = src[srow + 4 * (mx + 1)];
        *(uint32_t *)&cur[(L >> 2) * 16 + (L & 3) * 4] = cv;
        int g = lds_acq(&X.go);
        for (unsigned spins = 0; g <= mx; ++spins) {
            if (xr_give_up(spins, 1u << 24, ctl)) break;
            __builtin_amdgcn_s_sleep(1);
            g = lds_acq(&X.go);
        }
        const int p = mx & 1;
        // the screen: this wave's sub-mode r, then the four totals; every wave
        // publishes (a cancelled MB's totals are never used), so the exchange
        // completes whichever waves saw the cancel
        const uint8_t *top = (const uint8_t *)X.itop[p], *left = (const uint8_t *)X.ileft[p];
        const bool live = lds_acq(&X.cancel[p]) != mx + 1;
        Nb nbq = {};
        wave_sync();
        if (live) {
            nbq = xr_nb(top, left, X.icorner[p], my > 0, mx > 0, L);
            const int sr = xr_satd(cur, top, left, nbq, r, L);
            if (L == 0) X.sat[p][r] = sr;
        }
        wave_sync();
        if (L == 0) lds_rel(&X.sdone[r], mx + 1);
        int sv = L < 4 ? lds_acq(&X.sdone[L]) : mx + 1;
        for (unsigned spins = 0; !__all(sv > mx); ++spins) {
            if (xr_give_up(spins, 1u << 24, ctl)) break;
            __builtin_amdgcn_s_sleep(1);
            if (L < 4) sv = lds_acq(&X.sdone[L]);
        }
        if (live && lds_acq(&X.cancel[p]) != mx + 1) {
            int sel = -1, sbest = 0;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int v = X.sat[p][m];
                if (!nn_ok(m, nbq)) continue;
                if (sel < 0 || v < sbest) { sbest = v; sel = m; }
            }
            const int y = oy + (L >> 3), x = ox + (L & 7);
            X.ipred[y * 16 + x] = (uint8_t)ipred(sel, nbq, top, left, 16, x, y);
            wave_sync();
            region_code<0, true>(X.ws, r, cur + o, 16, X.ipred + o, 16, X.irec + o, 16, X.ilev, X.iflag, q.qp,
                                 q.lambda, &X.tab);
        }
        wave_sync();
        if (L == 0) lds_rel(&X.done[r], mx + 1);
    }
}

// The exact replay: one workgroup a MB row, taken by ticket (ctl[6]) in the
// order of `units` (row by row, every sequence's row my before any row my + 1),
// so the row above has always started.  Every wait is bounded (error flag).
extern "C" __global__ void __launch_bounds__(64 * XR_WAVES)
k_pxrep(const fdvg_seq *__restrict__ seqs, const int2 *__restrict__ units, int nunit, unsigned *ctl,
        const uint64_t *__restrict__ pcw, const uint64_t *__restrict__ edgew, fdvg_mbinfo *__restrict__ given,
        fdvg_pctx *__restrict__ pctx, uint64_t *__restrict__ mrec, uint64_t *__restrict__ botw,
        uint64_t *__restrict__ decw) {
    __shared__ XrLds X;
    __shared__ unsigned role_s;
    const int tid = threadIdx.x, wv = tid >> 6, L = tid & 63;
    if (tid == 0) {
        role_s = atomicAdd(ctl + 6, 1u);
        X.go = 0; X.cancel[0] = X.cancel[1] = 0;
    }
    if (tid < 4) { X.done[tid] = 0; X.sdone[tid] = 0; }
    tab8_fill(X.tab, tid);
    __syncthreads();
    const unsigned role = (unsigned)__builtin_amdgcn_readfirstlane((int)role_s);
    if (role >= (unsigned)nunit) return;
    const int2 u = units[role];
    const fdvg_seq q = seqs[u.x];
    if (wv == 0) xr_decide(X, q, pcw, edgew, ctl, u.y, mrec, botw, decw, given, pctx, L);
    else xr_trial(X, q, ctl, u.y, wv - 1, L);
}

// The intra MBs of the batch ordered by anti-diagonal (mx + my), all sequences
// interleaved: an intra MB depends only on its left, top and top-left
// neighbours, which lie on earlier diagonals, so every dependency comes earlier
// in the list and the MBs of one diagonal are independent.  ilist[0] = count,
// then the MB indices.  One workgroup: a histogram per diagonal, its exclusive
// scan, then the scatter (order within a diagonal is immaterial).  With more
// diagonals than the histogram holds, MB order (also dependency-safe).
#define PL_MAXD 2048
extern "C" __global__ void __launch_bounds__(1024)
k_plist(const fdvg_seq *__restrict__ seqs, int nseq, const fdvg_mbinfo *__restrict__ given, int total_mb,
        unsigned *__restrict__ ilist) {
    __shared__ unsigned hist[PL_MAXD];
    __shared__ int s_mb0[64], s_mw[64];
    __shared__ unsigned wsum[16];
    __shared__ unsigned base;
    __shared__ int maxd;
    const int t = threadIdx.x, L = t & 63, wv = t >> 6;
    const bool seq_lds = nseq <= 64;
    if (t < 64 && t < nseq) { s_mb0[t] = seqs[t].mb0; s_mw[t] = seqs[t].w >> 4; }
    if (t == 0) {
        int md = 0;
        for (int k = 0; k < nseq; ++k) md = max(md, (seqs[k].w >> 4) + (seqs[k].h >> 4) - 2);
        maxd = md;
        base = 0;
    }
    for (int d = t; d < PL_MAXD; d += blockDim.x) hist[d] = 0;
    __syncthreads();
    auto diag = [&](int mb) {
        int k = 0;
        if (seq_lds) { for (int kk = 1; kk < nseq; ++kk) if (mb >= s_mb0[kk]) k = kk; }
        else { for (int kk = 1; kk < nseq; ++kk) if (mb >= seqs[kk].mb0) k = kk; }
        const int mb0 = seq_lds ? s_mb0[k] : seqs[k].mb0, mw = seq_lds ? s_mw[k] : seqs[k].w >> 4;
        const int i = mb - mb0;
        return i % mw + i / mw;
    };
    if (maxd < PL_MAXD) {
        for (int mb = t; mb < total_mb; mb += blockDim.x)
            if (given[mb].mode == 2) atomicAdd(&hist[diag(mb)], 1u);
        __syncthreads();
        // exclusive scan of the histogram: two entries a thread, wave scans, then
        // the 16 wave totals
        {
            const unsigned a = hist[2 * t], b = hist[2 * t + 1];
            unsigned v = a + b;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned u = __shfl_up(v, o);
                if (L >= o) v += u;
            }
            if (L == 63) wsum[wv] = v;
            __syncthreads();
            unsigned off = 0;
            for (int k = 0; k < wv; ++k) off += wsum[k];
            const unsigned excl = off + v - (a + b);
            hist[2 * t] = excl;
            hist[2 * t + 1] = excl + a;
            if (t == 0) { unsigned tot = 0; for (int k = 0; k < 16; ++k) tot += wsum[k]; base = tot; }
        }
        __syncthreads();
        for (int mb = t; mb < total_mb; mb += blockDim.x)
            if (given[mb].mode == 2) ilist[1 + atomicAdd(&hist[diag(mb)], 1u)] = (unsigned)mb;
        if (t == 0) ilist[0] = base;
        return;
    }
    for (int c = 0; c < total_mb; c += 1024) {
        const int mb = c + t;
        const bool is = mb < total_mb && given[mb].mode == 2;
        const uint64_t bal = __ballot(is);
        const unsigned before = __builtin_popcountll(bal & ((1ull << L) - 1));
        if (L == 0) wsum[wv] = __builtin_popcountll(bal);
        __syncthreads();
        unsigned off = base;
        for (int k = 0; k < wv; ++k) off += wsum[k];
        if (is) ilist[1 + off + before] = (unsigned)mb;
        __syncthreads();
        if (t == 0) { unsigned tot = 0; for (int k = 0; k < 16; ++k) tot += wsum[k]; base += tot; }
        __syncthreads();
    }
    if (t == 0) ilist[0] = base;
}

__device__ __forceinline__ void ctx_from(MbCtx &c, const fdvg_pctx &X, int ht, int hl) {
    c.pmx = X.pmx; c.pmy = X.pmy; c.ncand = X.ncand; c.ht = ht; c.hl = hl;
#pragma unroll
    for (int k = 0; k < 3; ++k) { c.cx[k] = X.cx[k]; c.cy[k] = X.cy[k]; c.cr[k] = X.cr[k]; }
}

// Decided non-intra MBs, coded exactly (the given-decisions path of mb_decide).
extern "C" __global__ void __launch_bounds__(256)
k_pencode(const fdvg_seq *__restrict__ seqs, const fdvg_mbinfo *__restrict__ given,
          const fdvg_pctx *__restrict__ pctx, fdvg_mbrec *__restrict__ recs) {
    __shared__ Smem s;
    fdvg_seq q = seqs[blockIdx.y];
    int i;
    if (!xcd_mb(blockIdx.x, q.w >> 4, q.h >> 4, i)) return;
    const int mb = q.mb0 + i;
    const fdvg_mbinfo *gv = &given[mb];
    if (gv->mode == 2) return;
    q.given = given + q.mb0;
    const int mw = q.w >> 4, mx = i % mw, my = i / mw, bx = mx * 16, by = my * 16;
    mb_stage(s, q, mb, bx, by);
    MbCtx c;
    ctx_from(c, pctx[mb], my > 0, mx > 0);
    __syncthreads();
    int par = 0;
    MbDec D;
    mb_decide(s, q, c, gv, nullptr, par, D);
    mb_finish(s, q, mb, bx, by, c, D, &recs[mb]);
}

// The neighbours of an intra MB in the block-parallel pass (gather_nb,
// fdv.h:5201-5212): luma corner + top row into s.hbytes, left column into
// s.lcol[0..15], chroma corners + top rows into s.ctop and left columns into
// s.lcol[16..31].  Intra neighbours come from their handoff words, the others
// (coded by an earlier kernel) from the reconstruction planes; with gq NULL
// every neighbour comes from its handoff words.
__device__ __forceinline__ void nb_load_ih(Smem &s, const fdvg_seq &q, const fdvg_mbinfo *gq, const uint64_t *ih,
                                           unsigned *ctl, int mx, int my, int mw, int bx, int by) {
    const int tid = threadIdx.x, cbx = bx >> 1, cby = by >> 1, ht = my > 0, hl = mx > 0;
    const int i_top = (my - 1) * mw + mx, i_left = my * mw + mx - 1, i_tl = (my - 1) * mw + mx - 1;
    const bool top_i = ht && (!gq || gq[i_top].mode == 2), left_i = hl && (!gq || gq[i_left].mode == 2);
    const bool tl_i = ht && hl && (!gq || gq[i_tl].mode == 2);
    const uint64_t *ih0 = ih + (size_t)q.mb0 * IH_WORDS;
    const unsigned step = q.step;
    const gmem_t<uint8_t> *py = gmem(q.oy), *pu = gmem(q.ou), *pv = gmem(q.ov);
    const int ys = q.ys, cs = q.cs;
    if (tid < 16) {
        s.hbytes[1 + tid] = !ht ? 0 : top_i ? ih_byte(ih0 + (size_t)i_top * IH_WORDS, tid >> 2, tid & 3, step, ctl)
                                           : py[(ptrdiff_t)(by - 1) * ys + bx + tid];
    } else if (tid == 16) {
        s.hbytes[0] = !(ht && hl) ? 128 : tl_i ? ih_byte(ih0 + (size_t)i_tl * IH_WORDS, 3, 3, step, ctl)
                                               : py[(ptrdiff_t)(by - 1) * ys + bx - 1];
    } else if (tid >= 32 && tid < 48) {
        const int k = tid - 32;
        s.lcol[k] = !hl ? 0 : left_i ? ih_byte(ih0 + (size_t)i_left * IH_WORDS, 8 + (k >> 2), k & 3, step, ctl)
                                     : py[(ptrdiff_t)(by + k) * ys + bx - 1];
    } else if (tid >= 64 && tid < 82) {
        const int pl = (tid - 64) / 9, k = (tid - 64) % 9;
        const gmem_t<uint8_t> *p = pl ? pv : pu;
        if (k == 0)
            s.ctop[pl][0] = !(ht && hl) ? 128 : tl_i ? ih_byte(ih0 + (size_t)i_tl * IH_WORDS, pl ? 7 : 5, 3, step, ctl)
                                                     : p[(ptrdiff_t)(cby - 1) * cs + cbx - 1];
        else
            s.ctop[pl][k] = !ht ? 0 : top_i ? ih_byte(ih0 + (size_t)i_top * IH_WORDS, (pl ? 6 : 4) + ((k - 1) >> 2),
                                                      (k - 1) & 3, step, ctl)
                                            : p[(ptrdiff_t)(cby - 1) * cs + cbx + k - 1];
    } else if (tid >= 96 && tid < 112) {
        const int pl = (tid - 96) >> 3, k = (tid - 96) & 7;
        const gmem_t<uint8_t> *p = pl ? pv : pu;
        s.lcol[16 + 8 * pl + k] = !hl ? 0 : left_i ? ih_byte(ih0 + (size_t)i_left * IH_WORDS, (pl ? 14 : 12) + (k >> 2),
                                                             k & 3, step, ctl)
                                                   : p[(ptrdiff_t)(cby + k) * cs + cbx - 1];
    }
}

// The neighbours of an intra MB for a coding worker of the decisions, in two
// parts.  Luma (nb_load_rep): the top row, corner and left column of the
// decided reconstruction from the replays' edge words (BW_WORDS a MB), which
// never wait for a worker.  Chroma (nb_load_ihc): from the neighbours' handoff
// words, published by the workers that coded them.  The luma can then be coded
// as soon as the MB is decided; only the chroma waits for other workers.
__device__ __forceinline__ void nb_load_rep(Smem &s, const fdvg_seq &q, const uint64_t *botw, unsigned *ctl, int mx,
                                            int my, int mw) {
    const int tid = threadIdx.x, ht = my > 0, hl = mx > 0;
    const uint64_t *b0 = botw + (size_t)q.mb0 * BW_WORDS;
    const size_t i_top = (size_t)(my - 1) * mw + mx, i_left = (size_t)my * mw + mx - 1, i_tl = i_top - 1;
    if (tid < 16)
        s.hbytes[1 + tid] = !ht ? 0 : ih_byte(b0 + i_top * BW_WORDS, tid >> 2, tid & 3, q.step, ctl);
    else if (tid == 16)
        s.hbytes[0] = !(ht && hl) ? 128 : ih_byte(b0 + i_tl * BW_WORDS, 3, 3, q.step, ctl);
    else if (tid >= 32 && tid < 48)
        s.lcol[tid - 32] = !hl ? 0 : ih_byte(b0 + i_left * BW_WORDS, 4 + ((tid - 32) >> 2), tid & 3, q.step, ctl);
}
__device__ __forceinline__ void nb_load_ihc(Smem &s, const fdvg_seq &q, const uint64_t *ih, unsigned *ctl, int mx,
                                            int my, int mw) {
    const int tid = threadIdx.x, ht = my > 0, hl = mx > 0;
    const uint64_t *ih0 = ih + (size_t)q.mb0 * IH_WORDS;
    const size_t i_top = (size_t)(my - 1) * mw + mx, i_left = (size_t)my * mw + mx - 1, i_tl = i_top - 1;
    if (tid < 18) {
        const int pl = tid / 9, k = tid % 9;
        if (k == 0)
            s.ctop[pl][0] = !(ht && hl) ? 128 : ih_byte(ih0 + i_tl * IH_WORDS, pl ? 7 : 5, 3, q.step, ctl);
        else
            s.ctop[pl][k] = !ht ? 0 : ih_byte(ih0 + i_top * IH_WORDS, (pl ? 6 : 4) + ((k - 1) >> 2), (k - 1) & 3,
                                              q.step, ctl);
    } else if (tid >= 32 && tid < 48) {
        const int pl = (tid - 32) >> 3, k = (tid - 32) & 7;
        s.lcol[16 + 8 * pl + k] = !hl ? 0 : ih_byte(ih0 + i_left * IH_WORDS, (pl ? 14 : 12) + (k >> 2), k & 3,
                                                     q.step, ctl);
    }
}

// Decided intra MBs, in the order of ilist (every intra neighbour first): each
// waits for its left, top and top-left neighbours when those are intra, then
// reads the unfiltered reconstruction around it.
extern "C" __global__ void __launch_bounds__(256)
k_pintra(const fdvg_seq *__restrict__ seqs, int nseq, const fdvg_mbinfo *__restrict__ given,
         const fdvg_pctx *__restrict__ pctx, fdvg_mbrec *__restrict__ recs, unsigned *ctl,
         const unsigned *__restrict__ ilist, uint64_t *__restrict__ ih) {
    __shared__ Smem s;
    __shared__ int job_s;
    const int tid = threadIdx.x;
    const unsigned n = ilist[0];
    for (;;) {
        if (tid == 0) job_s = (int)atomicAdd(ctl + 4, 1u);
        __syncthreads();
        const unsigned job = (unsigned)__builtin_amdgcn_readfirstlane(job_s);   // uniform: scalar registers
        __syncthreads();
        if (job >= n) return;
        const int mb = __builtin_amdgcn_readfirstlane((int)ilist[1 + job]);
        int sq = 0;
        for (int k = 1; k < nseq; ++k) if (mb >= seqs[k].mb0) sq = k;
        fdvg_seq q = seqs[sq];
        q.given = given + q.mb0;
        const int i = mb - q.mb0, mw = q.w >> 4, mx = i % mw, my = i / mw, bx = mx * 16, by = my * 16;
        mb_stage(s, q, mb, bx, by);
        nb_load_ih(s, q, given + q.mb0, ih, ctl, mx, my, mw, bx, by);
        MbCtx c;
        ctx_from(c, pctx[mb], my > 0, mx > 0);
        __syncthreads();
        int par = 0;
        MbDec D;
        mb_decide(s, q, c, &given[mb], nullptr, par, D);
        const uint8_t *yrec = mb_finish(s, q, mb, bx, by, c, D, &recs[mb]);
        __syncthreads();                       // the chroma reconstruction is complete
        ih_publish(ih + (size_t)mb * IH_WORDS, yrec, s.crec, q.step);
    }
}

// The coding workers of the decisions (k_pdecide): the workers of XCD k take,
// by its own ticket (ctl[16 + k]), the MBs of its rows in the order in
// which the replays decide them (xorder, as k_pcosts), wait for the MB's
// decision words, code it as k_pencode / k_pintra do (staging only the
// reference window the decision reads), and publish its reconstruction's
// bottom row and right column as handoff words.  An intra MB waits for the
// handoff words of its left, top and top-left neighbours, which are earlier in
// the replays' order (in this XCD's list or another's).  Every wait is on a
// decision (made by a replay, which never waits for the workers) or on an
// earlier MB in that order, and is bounded (error flag on timeout); with every
// XCD holding workers, the earliest unfinished MB is always being coded.
extern "C" __global__ void __launch_bounds__(256)
k_pdecide(const fdvg_seq *__restrict__ seqs, int nseq, const unsigned *__restrict__ xorder,
          const int *__restrict__ xoff, unsigned *ctl, const uint64_t *__restrict__ decw,
          const uint64_t *__restrict__ botw, uint64_t *__restrict__ ih, fdvg_mbrec *__restrict__ recs) {
    __shared__ Smem s;
    __shared__ union {
        uint32_t w[PD_WORDS];
        struct { fdvg_mbinfo g; fdvg_pctx x; } d;
    } decb;
    const int tid = threadIdx.x;
    __shared__ int job_s;
    const int xk = blockIdx.x & 7, lo = xoff[xk], n = xoff[xk + 1] - lo;
    for (;;) {
        if (tid == 0) job_s = (int)atomicAdd(ctl + 16 + xk, 1u);
        __syncthreads();
        const int job = __builtin_amdgcn_readfirstlane(job_s);        // uniform: scalar registers
        __syncthreads();
        if (job >= n) return;
        const int mb = __builtin_amdgcn_readfirstlane((int)xorder[lo + job]);
        int sq = 0;
        for (int k = 1; k < nseq; ++k) if (mb >= seqs[k].mb0) sq = k;
        fdvg_seq q = seqs[sq];
        q.given = nullptr;
        const int i = mb - q.mb0, mw = q.w >> 4, mx = i % mw, my = i / mw, bx = mx * 16, by = my * 16;
        if (tid < PD_WORDS) {
            const uint64_t *w = decw + (size_t)mb * PD_WORDS + tid;
            uint64_t v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (unsigned spins = 0; (unsigned)(v >> 32) != q.step; ++spins) {
                if (spins > (1u << 21)) { atomicOr(ctl + 1, 1u); break; }
                __builtin_amdgcn_s_sleep(2);
                v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            decb.w[tid] = (uint32_t)v;
        }
        __syncthreads();
        const fdvg_mbinfo &G = decb.d.g;
        const fdvg_pctx &X = decb.d.x;
        const int gmode = G.mode;
        mb_stage(s, q, mb, bx, by, gmode, 1 << G.ref);
        if (gmode == 2) nb_load_rep(s, q, botw, ctl, mx, my, mw);
        MbCtx c;
        ctx_from(c, X, my > 0, mx > 0);
        __syncthreads();
        int par = 0;
        MbDec D;
        mb_decide(s, q, c, &G, nullptr, par, D);
        if (gmode == 2) {
            // the luma is coded; the chroma prediction waits for the neighbours' workers
            nb_load_ihc(s, q, ih, ctl, mx, my, mw);
            __syncthreads();
        }
        const uint8_t *yrec = mb_finish(s, q, mb, bx, by, c, D, &recs[mb]);
        __syncthreads();                       // the chroma reconstruction is complete
        ih_publish(ih + (size_t)mb * IH_WORDS, yrec, s.crec, q.step);
    }
}
static_assert(sizeof(fdvg_mbinfo) == 20 && sizeof(fdvg_pctx) == 20 && alignof(fdvg_pctx) <= 4,
              "decision words: 5 + 5 payload words");

// The block-parallel order of a batch: corder[] = every MB, ordered by
// mx + 2 my (then sequence, then row), the order in which the replays decide
// them (k_pcosts costs them in this order, the coding workers code them in it);
// and the replay units: units[] = (seq, row) for every MB row, ordered by row
// then sequence (a row's row above is in a lower-numbered unit).  Returns the
// unit count, or -1 on inconsistent sizes.
extern "C" int fdvg_pdecide_plan(const int *w, const int *h, const int *mb0, int nseq, int2 *units,
                                 unsigned *corder) {
    int nunit = 0, maxh = 0, maxt = 0, total = 0;
    for (int k = 0; k < nseq; ++k) {
        const int mw = w[k] >> 4, mh = h[k] >> 4;
        if (mw <= 0 || mh <= 0 || mb0[k] != total) return -1;
        if (mh > maxh) maxh = mh;
        if (mw - 1 + 2 * (mh - 1) > maxt) maxt = mw - 1 + 2 * (mh - 1);
        total += mw * mh;
    }
    for (int my = 0; my < maxh; ++my)
        for (int k = 0; k < nseq; ++k)
            if (my < (h[k] >> 4)) units[nunit++] = make_int2(k, my);
    int n = 0;
    for (int t = 0; t <= maxt; ++t)
        for (int k = 0; k < nseq; ++k) {
            const int mw = w[k] >> 4, mh = h[k] >> 4;
            for (int my = 0; my < mh; ++my) {
                const int mx = t - 2 * my;
                if (mx >= 0 && mx < mw) corder[n++] = (unsigned)(mb0[k] + my * mw + mx);
            }
        }
    return n == total ? nunit : -1;
}

// The per-XCD job lists of k_pcosts and k_pdecide: every MB of corder (the
// replays' order) goes, in that order, to the list of the XCD that owns its MB
// row, the rows of the batch (sequence after sequence) dealt to the 8 XCDs in
// turn.  An XCD then holds whole rows, whose neighbouring MBs share most of
// their reference windows, and every XCD has rows on the replays' front at
// any time (the front runs along the rows of every sequence at once).
// xorder holds the eight lists one after the other, xoff[k] .. xoff[k + 1]
// the k-th.  Returns the longest list's length.
extern "C" int fdvg_xcd_order(const int *w, const int *h, const int *mb0, int nseq, const unsigned *corder,
                              int total_mb, unsigned *xorder, int *xoff) {
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto xcd = [&](unsigned mb) {
        int k = 0, rows = 0;
        for (int j = 1; j < nseq; ++j) if ((int)mb >= mb0[j]) k = j;
        for (int j = 0; j < k; ++j) rows += h[j] >> 4;
        return (rows + ((int)mb - mb0[k]) / (w[k] >> 4)) & 7;
    };
    for (int t = 0; t < total_mb; ++t) ++cnt[xcd(corder[t])];
    xoff[0] = 0;
    int mx = 0;
    for (int k = 0; k < 8; ++k) { xoff[k + 1] = xoff[k] + cnt[k]; mx = cnt[k] > mx ? cnt[k] : mx; cnt[k] = 0; }
    for (int t = 0; t < total_mb; ++t) {
        const int k = xcd(corder[t]);
        xorder[xoff[k] + cnt[k]++] = corder[t];
    }
    return mx;
}

// Coding workers of k_pdecide: workers_per_cu a CU (at most what a CU holds);
// with workers_per_cu < 0, for a launch beside the nunit replays of k_pxrep:
// as many as leave, however they are placed (one a CU at worst), every
// replay not yet resident a CU of its XCD with no worker and no other replay;
// 0 when that leaves fewer than a quarter of the CUs.  Always a multiple of 8,
// so that every XCD holds workers.
extern "C" int fdvg_pdecide_grid(int nunit, int total_mb, int workers_per_cu) {
    static int ncu = 0, per = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return -1;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_pdecide, 256, 0) != hipSuccess) return -1;
    }
    if (workers_per_cu < 0) {
        // per XCD (workgroups go to the 8 XCDs round-robin): its CUs less the
        // replays it may hold
        const int wx = ncu / 8 - (nunit + 7) / 8;
        return wx < ncu / 32 ? 0 : 8 * wx;
    }
    const int wk = ncu * (workers_per_cu < per ? workers_per_cu : per);
    return wk < 8 ? 8 : wk / 8 * 8;
}

extern "C" int fdvg_launch_pdecide(const fdvg_seq *seqs, int nseq, const unsigned *xorder, const int *xoff, int grid,
                                   unsigned *ctl, const uint64_t *decw, const uint64_t *botw, uint64_t *ih,
                                   fdvg_mbrec *recs, hipStream_t st) {
    hipLaunchKernelGGL(k_pdecide, dim3(grid), dim3(256), 0, st, seqs, nseq, xorder, xoff, ctl, decw, botw, ih, recs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvg_launch_pxrep(const fdvg_seq *seqs, const int2 *units, int nunit, unsigned *ctl,
                                 const uint64_t *pcw, const uint64_t *edgew, fdvg_mbinfo *given, fdvg_pctx *pctx,
                                 uint64_t *mrec, uint64_t *botw, uint64_t *decw, hipStream_t st) {
    hipLaunchKernelGGL(k_pxrep, dim3(nunit), dim3(64 * XR_WAVES), 0, st, seqs, units, nunit, ctl, pcw, edgew, given,
                       pctx, mrec, botw, decw);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// The intra MBs of a batch in dependency order (k_plist), for other passes
// that code or decode intra MBs block-parallel.
extern "C" int fdvg_launch_plist(const fdvg_seq *seqs, int nseq, const fdvg_mbinfo *given, int total_mb,
                                 unsigned *ilist, hipStream_t st) {
    hipLaunchKernelGGL(k_plist, dim3(1), dim3(1024), 0, st, seqs, nseq, given, total_mb, ilist);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// xmax: the longest per-XCD list (fdvg_xcd_order)
extern "C" int fdvg_launch_pcosts(const fdvg_seq *seqs, int nseq, const unsigned *xorder, const int *xoff, int xmax,
                                  unsigned *ctl, uint64_t *pcw, uint64_t *edgew, hipStream_t st) {
    hipLaunchKernelGGL(k_pcosts, dim3(8 * xmax), dim3(256), 0, st, seqs, nseq, xorder, xoff, ctl, pcw, edgew);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvg_launch_pfinish(const fdvg_seq *seqs, int nseq, int xgrid, unsigned *ctl, fdvg_mbinfo *given,
                                   fdvg_pctx *pctx, unsigned *ilist, uint64_t *ih, int total_mb,
                                   fdvg_mbrec *recs, hipStream_t st) {
    static int cap_i = 0;
    if (!cap_i) {
        int dev = 0, ncu = 0, pi = 0;
        if (hipGetDevice(&dev) != hipSuccess) return -1;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pi, (const void *)k_pintra, 256, 0) != hipSuccess) return -1;
        cap_i = ncu * (pi > 0 ? pi : 1);
    }
    hipLaunchKernelGGL(k_plist, dim3(1), dim3(1024), 0, st, seqs, nseq, (const fdvg_mbinfo *)given, total_mb, ilist);
    hipLaunchKernelGGL(k_pencode, dim3(xgrid, nseq), dim3(256), 0, st, seqs, (const fdvg_mbinfo *)given,
                       (const fdvg_pctx *)pctx, recs);
    hipLaunchKernelGGL(k_pintra, dim3(total_mb < cap_i ? total_mb : cap_i), dim3(256), 0, st, seqs, nseq,
                       (const fdvg_mbinfo *)given, (const fdvg_pctx *)pctx, recs, ctl, (const unsigned *)ilist, ih);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// The decision stage as separate launches on one stream: the costs, the
// replays, the coding of the decisions.
extern "C" int fdvg_launch_parallel(const fdvg_seq *seqs, int nseq, int xgrid, const int2 *units, int nunit,
                                    const unsigned *xorder, const int *xoff, int xmax, unsigned *ctl, uint64_t *pcw,
                                    uint64_t *edgew, fdvg_mbinfo *given, fdvg_pctx *pctx, uint64_t *mrec,
                                    uint64_t *botw, uint64_t *decw, unsigned *ilist, uint64_t *ih, int total_mb,
                                    fdvg_mbrec *recs, hipStream_t st) {
    if (fdvg_launch_pcosts(seqs, nseq, xorder, xoff, xmax, ctl, pcw, edgew, st) != 0) return -1;
    if (fdvg_launch_pxrep(seqs, units, nunit, ctl, pcw, edgew, given, pctx, mrec, botw, decw, st) != 0) return -1;
    return fdvg_launch_pfinish(seqs, nseq, xgrid, ctl, given, pctx, ilist, ih, total_mb, recs, st);
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

// fdvgpu_intra.hip -- the key-frame (intra-only) plane codec on the device: the
// rate-distortion quadtree of fdv_image_encode (fdv.h:4845-4890; intra_walk
// 4746, intra_node 4494, intra_leaf4 3907, intra_leafn 4442, and code_residual
// 5300-5414 for the leaves above 4x4).
//
// k_intra_unit decides one 64x64 coding tree unit per workgroup.  Units are
// taken by ticket in the order of their wavefront key (column + 2 row, the
// planes of a batch interleaved).  A unit reads only its left neighbour's right
// column and the bottom rows of the three units above it, which the producers
// publish as words tagged with the call's step id: every unit waited for holds
// a smaller ticket and is running, and no fence is needed across XCDs.
//
// Inside a unit, wave 0 (the walker) walks the quadtree depth first as the
// reference recursion does -- a node's split trial (its children in z-order),
// then the node as one leaf, the leaf on ties -- and codes every 4x4 leaf
// itself.  A node's leaf trial reads only samples outside the node, so the
// walker posts it when it enters the node and the other waves (the coders)
// compute it while the walker codes the node's children: one job per (mode,
// 8x8 region), the deepest posted trial first, each job the region coder of the
// P-frame path.  A leaf's cost is an integer SSD plus lambda times integer bits,
// so the order in which the jobs finish does not change it.
//
// k_intra_emit then recodes the surviving leaves from the final reconstruction
// -- a surviving leaf's neighbours are the ones it was decided with -- into the
// level plane and the region flags, which the host serialises.
#include <hip/hip_runtime.h>
#include "fdvgpu_common.h"
#include "fdvgpu_resid.h"

#define IU 64                       // coding tree unit side (FDV_CTU, fdv.h:3651)
#define CVS 72                      // canvas row stride: columns -4 .. 67
#define NWI 8                       // waves of a unit workgroup: the walker and seven coders
#define NWE 4                       // waves of an emit workgroup
#define TREC_BYTES 17664
// canvas offset of unit-local sample (x, y), rows -1 .. 63
#define CV(y, x) ((((y) + 1) * CVS) + (x) + 4)

// trial reconstruction of mode m at depth d: n x n bytes (d 0: modes 0-2 only)
__host__ __device__ constexpr int trec_off(int d, int m) {
    return d == 0 ? m * 4096 : d == 1 ? 12288 + m * 1024 : d == 2 ? 16384 + m * 256 : 17408 + m * 64;
}
// Every sample of a 4x4 prediction is an entry of one array built from the
// edge E: A[i] = E[i] (copies), A[16 + i] = (E[i] + E[i+1] + 1) >> 1,
// A[32 + i] = (E[i] + 2 E[i+1] + E[i+2] + 2) >> 2, A[48] and A[49] the two
// corner taps of diagonal-down-left and horizontal-up, A[50] DC.  k_P4IDX
// holds, per (mode, row), the four entries' indices (the mapping p4 states).
static __device__ __constant__ uint32_t k_P4IDX[36] = {
    0x32323232, 0x32323232, 0x32323232, 0x32323232, 0x08070605, 0x08070605, 0x08070605, 0x08070605, 0x03030303,
    0x02020202, 0x01010101, 0x00000000, 0x28272625, 0x29282726, 0x2a292827, 0x302a2928, 0x26252423, 0x25242322,
    0x24232221, 0x23222120, 0x17161514, 0x26252423, 0x16151422, 0x25242321, 0x25242313, 0x23132212, 0x22122111,
    0x21112010, 0x18171615, 0x28272625, 0x19181716, 0x29282726, 0x20112112, 0x31102011, 0x00003110, 0x00000000};

// unit-local node index: depth-major, raster within a depth
__device__ __forceinline__ constexpr int nbase(int d) { return d == 0 ? 0 : d == 1 ? 1 : d == 2 ? 5 : 21; }

// One 4x4 candidate at each RDOQ truncation point t - 1 (t = 0 .. 16): the
// RDOQ's score, the cost of the point's own coding (clipped reconstruction
// SSD + lambda bits), its reconstruction; whether the RDOQ may stop there.
struct L4Res {
    double rdj[17], jf[17];
    uint32_t rw[17][4];
    uint8_t valid[17];
    int last;
};

struct __align__(16) IntraLds {
    uint8_t src[IU * IU];
    uint8_t cv[(IU + 1) * CVS];
    uint8_t trec[TREC_BYTES];
    uint8_t top[4][IU], left[4][IU];  // a posted trial's neighbours, per depth
    int prm[4][8];                    // n, ux, uy, mode mask, dc, plane a, b, c
    int gen[4], njobs[4], nreg[4], done[4];
    unsigned claim[4];                // generation << 16 | next job
    int accD[4][4], accB[4][4];       // per depth and mode: SSD and bits so far
    RegionWs ws[2];
    uint8_t hpred[NWI][64];
    int16_t hlev[NWI][64];
    uint8_t hflag[NWI];
    uint8_t pred4[9][16];             // the walker's nine 4x4 predictions
    int A[64];                        // the walker's 4x4 edge and its filtered taps
    uint32_t p4idx[36];               // k_P4IDX
    L4Res l4[9];                      // the walker's candidates at every truncation point
    uint8_t split[85], nmode[85];
    uint8_t cm[256];                  // the unit's 4x4 mode-context map
    int quit, plane, unit;
};

struct __align__(16) EmitLds {
    uint8_t src[IU * IU];
    uint8_t cv[(IU + 1) * CVS];
    RegionWs ws;
    uint8_t top[NWE][IU], left[NWE][IU];
    uint8_t hpred[NWE][64], hrec[NWE][64];
    int16_t hlev[NWE][64];
    uint8_t hflag[NWE];
    int E[NWE][16];
    double rdj[NWE][17];
    uint32_t rjob[64], qjob[256];      // surviving large-leaf regions, 4x4 leaves
    int nr, nq;
};

// ---------------------------------------------------------------- helpers

// z-order rank of a unit's 4x4 cell, vertical bit above horizontal (zidx4, fdv.h:3659)
__device__ __forceinline__ int zrank(int lx, int ly) {
    int z = 0;
#pragma unroll
    for (int b = 8; b; b >>= 1) z = (z << 2) | ((ly & b) ? 2 : 0) | ((lx & b) ? 1 : 0);
    return z;
}
// unit-local sample (x, y) is reconstructed while the 4x4 block at (ux, uy) is
// coded: a unit row above, or an earlier cell of this unit (cell_done, fdv.h:3677)
__device__ __forceinline__ bool cdone(int x, int y, int ux, int uy) {
    if (y < 0) return true;
    if (x >= IU) return false;
    return zrank(x >> 2, y >> 2) < zrank(ux >> 2, uy >> 2);
}
// the 4x4 modes the neighbours allow (mode_ok, fdv.h:3709)
__device__ __forceinline__ bool m4ok(int m, int ht, int htr, int hl, int htl) {
    return m == 0 ? true : m == 1 ? ht : (m == 2 || m == 8) ? hl : (m == 3 || m == 7) ? (ht && htr) : (ht && hl && htl);
}
// sample (x, y) of 4x4 mode m (H.264 8.3.1.2; fdv_intra_predict_4x4, fdv.h:2698)
// from the edge E (E[3 - i] left i, E[4] corner, E[5 + i] top i): every sample
// is a copy, a two-tap or a three-tap average of consecutive edge entries.
__device__ __forceinline__ int p4(const int *E, int m, int x, int y, int dcv) {
    int kind = 2, i = 0;
    switch (m) {
    case 1: kind = 0; i = 5 + x; break;
    case 2: kind = 0; i = 3 - y; break;
    case 3: if (x == 3 && y == 3) { kind = 3; i = 11; } else i = 5 + x + y; break;
    case 4: i = 3 + x - y; break;
    case 5: {
        const int z = 2 * x - y, k = x - (y >> 1);
        if (z >= 0) { kind = (z & 1) ? 2 : 1; i = (z & 1) ? 3 + k : 4 + k; }
        else i = z == -1 ? 3 : 4 - y;
        break;
    }
    case 6: {
        const int z = 2 * y - x, k = y - (x >> 1);
        if (z >= 0) { kind = (z & 1) ? 2 : 1; i = 3 - k; }
        else i = z == -1 ? 3 : 2 + x;
        break;
    }
    case 7: kind = (y & 1) ? 2 : 1; i = 5 + x + (y >> 1); break;
    case 8: {
        const int z = x + 2 * y, k = y + (x >> 1);
        if (z < 5) { kind = (z & 1) ? 2 : 1; i = (z & 1) ? 1 - k : 2 - k; }
        else if (z == 5) kind = 4;
        else kind = 0;
        break;
    }
    default: return dcv;
    }
    const int a = E[i], b = E[i + 1], c = E[i + 2];
    return kind == 0 ? a : kind == 1 ? (a + b + 1) >> 1 : kind == 2 ? (a + 2 * b + c + 2) >> 2 :
           kind == 3 ? (a + 3 * b + 2) >> 2 : (3 * a + b + 2) >> 2;
}

// The edge of the 4x4 block at unit-local (ux, uy) from a canvas, into E (all
// lanes write the same values; lanes 0..15 do the writing); returns the DC
// value and the availability flags (gather_neighbors, fdv.h:3689).
__device__ __forceinline__ int edge4(const uint8_t *cv, int *E, int ux, int uy, int bx, int by, int w,
                                     int &ht, int &hl, int &htr, int &htl, int L) {
    ht = by > 0; hl = bx > 0; htl = ht && hl;
    htr = ht && bx + 4 < w && cdone(ux + 4, uy - 1, ux, uy);
    int v = 0;
    if (L < 4) v = hl ? cv[CV(uy + L, ux - 1)] : 0;                       // E[3 - L]
    else if (L == 4) v = htl ? cv[CV(uy - 1, ux - 1)] : 128;
    else if (L < 13) {
        const int j = L - 5;
        v = ht ? cv[CV(uy - 1, ux + (j < 4 || htr ? j : 3))] : 0;
    }
    if (L < 16) E[L < 4 ? 3 - L : L] = v;
    // DC (fdv.h:2806-2822)
    const int tv = L >= 5 && L < 9 ? v : 0, lv = L < 4 ? v : 0;
    const int st = row_sum16(tv), sl = row_sum16(lv);
    const int s_t = __builtin_amdgcn_readlane(st, 0), s_l = __builtin_amdgcn_readlane(sl, 0);
    return ht && hl ? (s_t + s_l + 4) >> 3 : ht ? (s_t + 2) >> 2 : hl ? (s_l + 2) >> 2 : 128;
}

// The final coding of a 4x4 leaf with RDOQ (the second pass of intra_leaf4,
// fdv.h:4067-4096): lanes 0..15, every lane holding the block.  Lane c scores
// the RDOQ truncation point c - 1 (lane 0 also 15); lane 0 applies the
// reference's selection, reconstructs into rw[4] and returns the levels in
// lf[16], the SSD and the bits (mode symbol included).  pw / sw: the block's
// prediction and source rows as words.
__device__ __forceinline__ void final4(const uint32_t pw[4], const uint32_t sw[4], int qp, double lambda, double *rdj,
                                       int L, uint32_t rw[4], int lf[16], int &e, int &bits) {
    const int m6 = qp % 6, sh = 1 << (qp / 6), qb = 15 + qp / 6, off = (1 << qb) / 3;
    const int mf0 = k_MF[m6][0], mf1 = k_MF[m6][1], mf2 = k_MF[m6][2];
    const int dv0 = k_DV[m6][0], dv1 = k_DV[m6][1], dv2 = k_DV[m6][2];
    int cu[16], pr[16], res[16], lv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        cu[i] = (sw[i >> 2] >> (8 * (i & 3))) & 0xff;
        pr[i] = (pw[i >> 2] >> (8 * (i & 3))) & 0xff;
        res[i] = cu[i] - pr[i];
    }
    {
        int co[16];
        fdct4_reg(res, co);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int mf = cls4(i) == 0 ? mf0 : cls4(i) == 1 ? mf1 : mf2;
            const int a = co[i] < 0 ? -co[i] : co[i];
            const int v = (a * mf + off) >> qb;
            lv[i] = co[i] < 0 ? -v : v;
        }
    }
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
    if (L < 16) {
        for (int Lc = L - 1; Lc <= last; Lc += 16) {
            int b = bits0;
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k > Lc && k <= last) b -= drop[k];
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
            rdj[Lc + 1] = (double)D + lambda * (double)b;
        }
    }
    wave_sync();
    e = 0; bits = 0;
    if (L == 0) {
        double bestJ = -1.0;
        int bestL = last;
#pragma unroll
        for (int k = 16; k >= 0; --k) {
            const int Lc = k - 1;
            const bool valid = Lc <= last && (Lc < 0 || ((nzm >> Lc) & 1));
            if (valid) {
                const double J = rdj[k];
                if (bestJ < 0.0 || J < bestJ) { bestJ = J; bestL = Lc; }
            }
        }
        int dq[16], rr[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            lf[i] = k_INV4c(i) <= bestL ? lv[i] : 0;
            const int dv = cls4(i) == 0 ? dv0 : cls4(i) == 1 ? dv1 : dv2;
            dq[i] = (lf[i] * dv) * sh;
        }
        idct4_reg(dq, rr);
#pragma unroll
        for (int y = 0; y < 4; ++y) rw[y] = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int v = clip8(pr[i] + rr[i]);
            rw[i >> 2] |= (uint32_t)v << (8 * (i & 3));
            const int d = cu[i] - v;
            e += d * d;
        }
        int lastF = -1;
#pragma unroll
        for (int k = 0; k < 16; ++k) if (lf[k_SCAN4c(k)]) lastF = k;
        bits = 3 + egbits((uint32_t)(lastF + 1));                       // FDV_BITS_MODE (fdv.h:3351)
#pragma unroll
        for (int k = 0; k < 16; ++k) if (k <= lastF) bits += egbits(zmap(lf[k_SCAN4c(k)]));
    }
    wave_sync();
}

// A large leaf's neighbours and prediction parameters (gather_nb, fdv.h:5201;
// fdv_intra_nxn_inner 2879, plane_pred 2851): top / left n samples into the
// given buffers, the allowed-mode mask, DC, and the plane fit's a, b, c.
__device__ __forceinline__ void leafn_setup(const uint8_t *cv, int ux, int uy, int n, int bx, int by,
                                            uint8_t *top, uint8_t *left, int prm[8], int L) {
    const int ht = by > 0, hl = bx > 0;
    const int tv = L < n && ht ? cv[CV(uy - 1, ux + L)] : 0;
    const int lv = L < n && hl ? cv[CV(uy + L, ux - 1)] : 0;
    if (L < n) { top[L] = (uint8_t)tv; left[L] = (uint8_t)lv; }
    const int tl = ht && hl ? cv[CV(uy - 1, ux - 1)] : 128;
    const int st = wave_sum(tv), sl = wave_sum(lv);
    const int dc = ht && hl ? (st + sl + n) / (2 * n) : ht ? (st + n / 2) / n : hl ? (sl + n / 2) / n : 128;
    int mask = 1 | (ht ? 2 : 0) | (hl ? 4 : 0) | (ht && hl && n <= 32 ? 8 : 0);
    int a = 0, b = 0, c = 0;
    if (mask & 8) {
        const int hn = n / 2;
        wave_sync();
        int hterm = 0, vterm = 0;
        if (L < hn) {
            const int k = L + 1, j = hn - 2 - L;
            hterm = k * ((int)top[hn - 1 + k] - (j < 0 ? tl : (int)top[j]));
            vterm = k * ((int)left[hn - 1 + k] - (j < 0 ? tl : (int)left[j]));
        }
        const int H = wave_sum(hterm), V = wave_sum(vterm);
        a = 16 * ((int)left[n - 1] + (int)top[n - 1]);
        if (n == 32)      { b = (11 * H + 512) >> 10; c = (11 * V + 512) >> 10; }
        else if (n == 16) { b = (5 * H + 32) >> 6;    c = (5 * V + 32) >> 6; }
        else              { b = (17 * H + 16) >> 5;   c = (17 * V + 16) >> 5; }
    }
    prm[0] = n; prm[1] = ux; prm[2] = uy; prm[3] = mask; prm[4] = dc; prm[5] = a; prm[6] = b; prm[7] = c;
}
// sample (px, py) of a large leaf's mode m (node coordinates)
__device__ __forceinline__ int pn(const int prm[8], const uint8_t *top, const uint8_t *left, int m, int px, int py) {
    const int n = prm[0], c0 = n / 2 - 1;
    if (m == 1) return top[px];
    if (m == 2) return left[py];
    if (m == 3) return clip8((prm[5] + prm[7] * (py - c0) + 16 + prm[6] * (px - c0)) >> 5);
    return prm[4];
}
// wave-uniform minimum of unsigned values
__device__ __forceinline__ unsigned wave_umin(unsigned v) {
    v = min(v, (unsigned)dpp_mov<DPP_XOR1>((int)v)); v = min(v, (unsigned)dpp_mov<DPP_XOR2>((int)v));
    v = min(v, (unsigned)dpp_mov<DPP_ROR4>((int)v)); v = min(v, (unsigned)dpp_mov<DPP_ROR8>((int)v));
    return min(min((unsigned)__builtin_amdgcn_readlane((int)v, 0), (unsigned)__builtin_amdgcn_readlane((int)v, 16)),
               min((unsigned)__builtin_amdgcn_readlane((int)v, 32), (unsigned)__builtin_amdgcn_readlane((int)v, 48)));
}
// a large leaf's mode in the 4x4 mode context (nn_to_4x4_mode, fdv.h:4400)
__device__ __forceinline__ int nn_ctx(int m) { return m == 1 ? 1 : m == 2 ? 2 : 0; }

// ------------------------------------------------------------ the walker

struct UnitCtx {
    int x0, y0, w, h, qp;
    double lambda;
};

// A 4x4 leaf (intra_leaf4, fdv.h:3907-4108) by the walker wave: the SATD screen
// over the allowed modes, the first ntry by SATD (selection with exchanges, as
// the reference orders them) coded without RDOQ and compared in D + lambda*R,
// pruned once a candidate's SATD passes 5/4 of the first's; the winner recoded
// with RDOQ into the canvas.  Returns the winner's cost (wave-uniform).
__device__ __forceinline__ double leaf4(IntraLds &S, const UnitCtx &U, int ux, int uy) {
    const int L = threadIdx.x & 63;
    const int bx = U.x0 + ux, by = U.y0 + uy;
    int ht, hl, htr, htl;
    const int dcv = edge4(S.cv, S.A, ux, uy, bx, by, U.w, ht, hl, htr, htl, L);
    wave_sync();
    {
        int a = 0;
        if (L >= 16 && L < 28) a = (S.A[L - 16] + S.A[L - 15] + 1) >> 1;
        else if (L >= 32 && L < 43) a = (S.A[L - 32] + 2 * S.A[L - 31] + S.A[L - 30] + 2) >> 2;
        else if (L == 48) a = (S.A[11] + 3 * S.A[12] + 2) >> 2;
        else if (L == 49) a = (3 * S.A[0] + S.A[1] + 2) >> 2;
        else if (L == 50) a = dcv;
        if (L >= 16) S.A[L] = a;
    }
    wave_sync();
    // nine predictions and their SATD: lane 4m + y takes row y of mode m
    const int m = L >> 2, y = L & 3;
    const bool act = L < 36 && m4ok(m, ht, htr, hl, htl);
    int h[4];
    {
        const uint32_t sw = *(const uint32_t *)&S.src[(uy + y) * IU + ux];
        const uint32_t iw = L < 36 ? S.p4idx[L] : 0u;
        int d[4];
        uint32_t pw = 0;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            const int p = act ? S.A[(iw >> (8 * x)) & 0xff] : 0;
            pw |= (uint32_t)p << (8 * x);
            d[x] = (int)((sw >> (8 * x)) & 0xff) - p;
        }
        if (L < 36) *(uint32_t *)&S.pred4[m][y * 4] = pw;
        const int a0 = d[0] + d[1], a1 = d[0] - d[1], a2 = d[2] + d[3], a3 = d[2] - d[3];
        h[0] = a0 + a2; h[1] = a1 + a3; h[2] = a0 - a2; h[3] = a1 - a3;
    }
    int sacc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int v = h[j], p = dpp_mov<DPP_XOR1>(v);
        const int s = (y & 1) ? p - v : v - p, a = v + p;
        const int xv = (y & 1) ? s : a, x2 = dpp_mov<DPP_XOR2>(xv);
        const int o = (y & 2) ? x2 - xv : xv + x2;
        sacc += o < 0 ? -o : o;
    }
    sacc += dpp_mov<DPP_XOR1>(sacc);
    sacc += dpp_mov<DPP_XOR2>(sacc);
    wave_sync();                                      // the predictions, for the candidate lanes
    // candidate list in mode order (position = rank among the allowed modes),
    // then the reference's ordering of the first ntry: selection with
    // exchanges (fdv.h:3956-3969)
    int ks[9], km[9];
#pragma unroll
    for (int mm = 0; mm < 9; ++mm) { ks[mm] = 0x7fffffff; km[mm] = 0; }
    const bool rep_l = (L & 3) == 0 && act;
    const uint64_t okb = __ballot(rep_l);
    int okmask = 0;
#pragma unroll
    for (int mm = 0; mm < 9; ++mm) okmask |= (int)((okb >> (4 * mm)) & 1) << mm;
    const int nc = __builtin_popcount(okmask);
    const int ntry = U.qp < 12 ? nc : (nc < 2 ? nc : 2);
    if (ntry <= 2) {
        // two rounds as minima of (SATD, position) keys: the first pick is the
        // lowest; the exchange puts the first position's entry where the first
        // pick was, and the second pick is the lowest of the rest by the new
        // positions
        const int slot = __builtin_popcount(okmask & ((1 << m) - 1));
        const unsigned k0 = wave_umin(rep_l ? (unsigned)sacc << 4 | (unsigned)slot : 0xffffffffu);
        const int i0 = (int)(k0 & 15);
        const uint64_t b0 = __ballot(rep_l && slot == i0), bz = __ballot(rep_l && slot == 0);
        ks[0] = (int)(k0 >> 4); km[0] = (int)(__builtin_ctzll(b0) >> 2);
        if (ntry == 2) {
            const unsigned k1 = wave_umin(rep_l && slot != i0 ? (unsigned)sacc << 4 | (unsigned)(slot == 0 ? i0 : slot)
                                                              : 0xffffffffu);
            const int p1 = (int)(k1 & 15);
            const uint64_t b1 = p1 == i0 ? bz : __ballot(rep_l && slot == p1);
            ks[1] = (int)(k1 >> 4); km[1] = (int)(__builtin_ctzll(b1) >> 2);
        }
    } else {
        int n = 0;
#pragma unroll
        for (int mm = 0; mm < 9; ++mm) {
            const int sv = __builtin_amdgcn_readlane(sacc, 4 * mm);
            if ((okmask >> mm) & 1) {
#pragma unroll
                for (int s2 = 0; s2 < 9; ++s2) if (s2 == n) { ks[s2] = sv; km[s2] = mm; }
                ++n;
            }
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) {
        if (a < ntry) {
            int bst = a, bv = ks[a];
#pragma unroll
            for (int b = a + 1; b < 9; ++b) if (b < nc && ks[b] < bv) { bv = ks[b]; bst = b; }
#pragma unroll
            for (int b = a + 1; b < 9; ++b)
                if (b == bst) { const int t = ks[a]; ks[a] = ks[b]; ks[b] = t; const int u = km[a]; km[a] = km[b]; km[b] = u; }
        }
        }
    }
    int ne = ntry;
    if (U.qp >= 12) {
#pragma unroll
        for (int c = 8; c >= 1; --c)
            if (c < ntry && (long long)ks[c] * 4 > (long long)ks[0] * 5) ne = c;
    }
    // One pass per three candidates: lane 21 g + t takes candidate 3 pass + g
    // at RDOQ truncation point t - 1 (t = 0 .. 16), every lane holding the
    // candidate's block.  Each lane scores its truncation point as the RDOQ
    // does (unclipped residual SSD, the RDOQ's rate) and also codes it: the
    // clipped reconstruction, its SSD and its bits.  The point with no
    // truncation is the candidate's own coding without RDOQ, so the same pass
    // gives the candidates' comparison (the first pass of intra_leaf4) and,
    // for the winner, the RDOQ's choice and its result.
    const uint32_t *srow = (const uint32_t *)&S.src[uy * IU + ux];
    const uint32_t sw[4] = {srow[0], srow[IU / 4], srow[2 * IU / 4], srow[3 * IU / 4]};
    {
        const int qp = U.qp;
        const int m6 = qp % 6, sh = 1 << (qp / 6), qb = 15 + qp / 6, off = (1 << qb) / 3;
        const int mf0 = k_MF[m6][0], mf1 = k_MF[m6][1], mf2 = k_MF[m6][2];
        const int dv0 = k_DV[m6][0], dv1 = k_DV[m6][1], dv2 = k_DV[m6][2];
        const int grp = L / 21, t = L - 21 * grp, Lc = t - 1;
        for (int pass = 0; pass * 3 < ne; ++pass) {
            const int c = pass * 3 + grp;
            int mc = km[0];
#pragma unroll
            for (int cc = 1; cc < 9; ++cc) if (c == cc) mc = km[cc];
            if (grp < 3 && t <= 16 && c < ne) {
                const uint32_t *pp = (const uint32_t *)S.pred4[mc];
                int cu[16], pr[16], res[16], lv[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    cu[i] = (sw[i >> 2] >> (8 * (i & 3))) & 0xff;
                    pr[i] = (pp[i >> 2] >> (8 * (i & 3))) & 0xff;
                    res[i] = cu[i] - pr[i];
                }
                {
                    int co[16];
                    fdct4_reg(res, co);
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const int mf = cls4(i) == 0 ? mf0 : cls4(i) == 1 ? mf1 : mf2;
                        const int a = co[i] < 0 ? -co[i] : co[i];
                        const int v = (a * mf + off) >> qb;
                        lv[i] = co[i] < 0 ? -v : v;
                    }
                }
                int last = -1, nzm = 0, drop[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const uint32_t z = zmap(lv[k_SCAN4c(k)]);
                    drop[k] = 8 * leblen(z);
                    if (z) { last = k; nzm |= 1 << k; }
                }
                L4Res &R = S.l4[c];
                if (t == 0) R.last = last;
                const bool valid = Lc <= last && (Lc < 0 || ((nzm >> Lc) & 1));
                R.valid[t] = (uint8_t)valid;
                if (Lc <= last) {
                    int bits0 = egbits((uint32_t)(last + 1));
#pragma unroll
                    for (int k = 0; k < 16; ++k) if (k <= last) bits0 += egbits(zmap(lv[k_SCAN4c(k)]));
                    int rb = bits0, lt[16], dq[16], rr[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) if (k > Lc && k <= last) rb -= drop[k];
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        lt[i] = k_INV4c(i) <= Lc ? lv[i] : 0;
                        const int dv = cls4(i) == 0 ? dv0 : cls4(i) == 1 ? dv1 : dv2;
                        dq[i] = (lt[i] * dv) * sh;
                    }
                    idct4_reg(dq, rr);
                    int D = 0, e = 0;
                    uint32_t rw[4] = {0, 0, 0, 0};
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const int d = res[i] - rr[i];
                        D += d * d;
                        const int v = clip8(pr[i] + rr[i]);
                        rw[i >> 2] |= (uint32_t)v << (8 * (i & 3));
                        const int d2 = cu[i] - v;
                        e += d2 * d2;
                    }
                    int lastF = -1;
#pragma unroll
                    for (int k = 0; k < 16; ++k) if (lt[k_SCAN4c(k)]) lastF = k;
                    int fb = 3 + egbits((uint32_t)(lastF + 1));               // FDV_BITS_MODE (fdv.h:3351)
#pragma unroll
                    for (int k = 0; k < 16; ++k) if (k <= lastF) fb += egbits(zmap(lt[k_SCAN4c(k)]));
                    R.rdj[t] = (double)D + U.lambda * (double)rb;
                    R.jf[t] = (double)e + U.lambda * (double)fb;
#pragma unroll
                    for (int r = 0; r < 4; ++r) R.rw[t][r] = rw[r];
                }
            }
        }
    }
    wave_sync();
    // the candidates' costs without RDOQ, in the reference's order of
    // comparison: lane c reads candidate c's untruncated point
    int bc = 0;
    {
        double Jc = 0.0;
        if (L < ne) Jc = S.l4[L].jf[S.l4[L].last + 1];
        const uint64_t jb = __double_as_longlong(Jc);
        const uint32_t jlo = (uint32_t)jb, jhi = (uint32_t)(jb >> 32);
        double best = -1.0;
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            if (c < ne) {
                const uint64_t v = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(jhi, c) << 32 |
                                   (uint32_t)__builtin_amdgcn_readlane(jlo, c);
                const double J = __longlong_as_double((long long)v);
                if (best < 0.0 || J < best) { best = J; bc = c; }
            }
        }
    }
    int mw = km[0];
#pragma unroll
    for (int c = 1; c < 9; ++c) if (bc == c) mw = km[c];
    // the RDOQ's choice for the winner: its scan from the last level down,
    // the first lowest score among the points it may stop at
    int bt = 0;
    {
        const L4Res &R = S.l4[bc];
        const int last = R.last;
        double rj = 0.0;
        int vl = 0;
        if (L <= last + 1) { rj = R.rdj[L]; vl = R.valid[L]; }
        const uint64_t vmask = __ballot(vl != 0);
        const uint64_t jb = __double_as_longlong(rj);
        const uint32_t jlo = (uint32_t)jb, jhi = (uint32_t)(jb >> 32);
        double bestJ = -1.0;
        bt = last + 1;
#pragma unroll
        for (int k = 16; k >= 0; --k) {
            if (k <= last + 1 && ((vmask >> k) & 1)) {
                const uint64_t v = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(jhi, k) << 32 |
                                   (uint32_t)__builtin_amdgcn_readlane(jlo, k);
                const double J = __longlong_as_double((long long)v);
                if (bestJ < 0.0 || J < bestJ) { bestJ = J; bt = k; }
            }
        }
    }
    const L4Res &W = S.l4[bc];
    if (L < 4) *(uint32_t *)&S.cv[CV(uy + L, ux)] = W.rw[bt][L];
    if (L == 0) S.cm[(uy >> 2) * 16 + (ux >> 2)] = (uint8_t)mw;
    const double J = W.jf[bt];
    wave_sync();
    const uint64_t jb = __double_as_longlong(J);
    const uint64_t v = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((uint32_t)(jb >> 32), 0) << 32 |
                       (uint32_t)__builtin_amdgcn_readlane((uint32_t)jb, 0);
    return __longlong_as_double((long long)v);
}

// Posts the leaf trial of the depth-D node at (ux, uy) for the coders.
template <int D>
__device__ __forceinline__ void post(IntraLds &S, const UnitCtx &U, int ux, int uy) {
    constexpr int n = IU >> D;
    const int L = threadIdx.x & 63;
    int prm[8];
    leafn_setup(S.cv, ux, uy, n, U.x0 + ux, U.y0 + uy, S.top[D], S.left[D], prm, L);
    if (L < 4) { S.accD[D][L] = 0; S.accB[D][L] = 0; }
    if (L == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) S.prm[D][k] = prm[k];
        const int nreg = (n / 8) * (n / 8);
        S.nreg[D] = nreg;
        S.njobs[D] = __builtin_popcount(prm[3]) * nreg;
        S.done[D] = 0;
        S.gen[D] = S.gen[D] + 1;
    }
    wave_sync();
    if (L == 0)
        __hip_atomic_exchange(&S.claim[D], (unsigned)S.gen[D] << 16, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One quadtree node at depth D (intra_node, fdv.h:4494-4584): a split implied
// by the plane's edge, or the split trial against the posted leaf trial, the
// leaf on ties.  Returns the winner's cost; the canvas holds the winner.
template <int D>
__device__ double walk(IntraLds &S, const UnitCtx &U, int ux, int uy, unsigned *err) {
    constexpr int n = IU >> D, hn = n / 2;
    const int L = threadIdx.x & 63;
    if constexpr (n == 4) {
        return leaf4(S, U, ux, uy);
    } else {
        const int li = nbase(D) + (uy / n) * (1 << D) + ux / n;
        const bool implied = U.x0 + ux + n > U.w || U.y0 + uy + n > U.h;
        if (!implied) post<D>(S, U, ux, uy);
        double J = 0.0;
#pragma unroll 1
        for (int k = 0; k < 4; ++k) {
            const int cx = ux + (k & 1) * hn, cy = uy + (k >> 1) * hn;
            if (U.x0 + cx < U.w && U.y0 + cy < U.h) J += walk<D + 1>(S, U, cx, cy, err);
        }
        if (implied) {
            if (L == 0) S.split[li] = 2;
            return J;
        }
        // the leaf trial's costs
        const int nj = S.njobs[D];
        for (unsigned spins = 0; __hip_atomic_load(&S.done[D], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < nj; ++spins) {
            if (spins > (1u << 24)) { if (L == 0) atomicOr(err, 2u); break; }
            __builtin_amdgcn_s_sleep(1);
        }
        const int mask = S.prm[D][3];
        double Jl = -1.0;
        int bm = 0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (!((mask >> m) & 1)) continue;
            const double Jm = (double)S.accD[D][m] + U.lambda * (double)(3 + S.accB[D][m]);
            if (Jl < 0.0 || Jm < Jl) { Jl = Jm; bm = m; }
        }
        const bool split = J < Jl;
        if (L == 0) { S.split[li] = split ? 1 : 0; S.nmode[li] = (uint8_t)bm; }
        if (!split) {
            const uint8_t *tr = S.trec + trec_off(D, bm);
            for (int i = L; i < n * n / 4; i += 64) {
                const int yy = (4 * i) / n, xx = (4 * i) % n;
                *(uint32_t *)&S.cv[CV(uy + yy, ux + xx)] = *(const uint32_t *)&tr[4 * i];
            }
            const int cells = n / 4, v = nn_ctx(bm);
            for (int i = L; i < cells * cells; i += 64) S.cm[((uy >> 2) + i / cells) * 16 + (ux >> 2) + i % cells] = (uint8_t)v;
        }
        wave_sync();
        return split ? J : Jl;
    }
}

// ------------------------------------------------------------- the coders

// One job of a posted trial: the 8x8 region r of mode m at depth d.
__device__ __forceinline__ void coder_job(IntraLds &S, const UnitCtx &U, int d, int j, int slot, int L) {
    const int *prm = S.prm[d];
    const int n = prm[0], ux = prm[1], uy = prm[2], mask = prm[3];
    const int nreg = S.nreg[d], mi = j / nreg, r = j % nreg, nr = n / 8;
    int m = 0;
    for (int k = 0, seen = 0; k < 4; ++k)
        if ((mask >> k) & 1) { if (seen == mi) { m = k; break; } ++seen; }
    const int rx = (r % nr) * 8, ry = (r / nr) * 8;
    S.hpred[slot][L] = (uint8_t)pn(prm, S.top[d], S.left[d], m, rx + (L & 7), ry + (L >> 3));
    wave_sync();
    const int g = slot >> 2, rs = slot & 3;
    region_code<0>(S.ws[g], rs, &S.src[(uy + ry) * IU + ux + rx], IU, S.hpred[slot], 8,
                   S.trec + trec_off(d, m) + ry * n + rx, n, &S.hlev[g * 4][0], &S.hflag[g * 4], U.qp, U.lambda);
    wave_sync();
    if (L == 0) {
        atomicAdd(&S.accD[d][m], S.ws[g].e[rs]);
        atomicAdd(&S.accB[d][m], S.ws[g].b[rs]);
        __hip_atomic_fetch_add(&S.done[d], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

__device__ __forceinline__ void coder_loop(IntraLds &S, const UnitCtx &U, int slot, unsigned *err) {
    const int L = threadIdx.x & 63;
    for (unsigned idle = 0;;) {
        int d = -1, j = 0;
        for (int dd = 3; dd >= 0 && d < 0; --dd) {
            const unsigned v = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&S.claim[dd], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
            const int g = __builtin_amdgcn_readfirstlane(S.gen[dd]), nj = __builtin_amdgcn_readfirstlane(S.njobs[dd]);
            if ((int)(v >> 16) != g || (int)(v & 0xffff) >= nj) continue;
            unsigned got = 0;
            if (L == 0) got = __hip_atomic_fetch_add(&S.claim[dd], 1u, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            got = __builtin_amdgcn_readfirstlane(got);
            const int g2 = __builtin_amdgcn_readfirstlane(S.gen[dd]), nj2 = __builtin_amdgcn_readfirstlane(S.njobs[dd]);
            if ((int)(got >> 16) == g2 && (int)(got & 0xffff) < nj2) { d = dd; j = (int)(got & 0xffff); }
        }
        if (d >= 0) { coder_job(S, U, d, j, slot, L); idle = 0; continue; }
        if (__hip_atomic_load(&S.quit, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
        if (++idle > (1u << 26)) { if (L == 0) atomicOr(err, 4u); break; }
        __builtin_amdgcn_s_sleep(1);
    }
}

// ---------------------------------------------------------------- kernels

// ctl[0]: ticket counter (zero before the launch); ctl[1]: error flags.
extern "C" __global__ void __launch_bounds__(64 * NWI)
k_intra_unit(const fdvg_iplane *__restrict__ planes, const uint32_t *__restrict__ order, int nunits, unsigned *ctl,
             unsigned step) {
    __shared__ IntraLds S;
    const int tid = threadIdx.x, wv = tid >> 6, L = tid & 63;
    if (tid == 0) {
        const unsigned t = atomicAdd(ctl, 1u);
        const uint32_t o = t < (unsigned)nunits ? order[t] : 0xffffffffu;
        S.plane = (int)(o >> 20); S.unit = (int)(o & 0xfffff);
    }
    for (int i = tid; i < 4; i += blockDim.x) { S.claim[i] = 0; S.gen[i] = 0; S.njobs[i] = 0; S.done[i] = 0; }
    for (int i = tid; i < 85; i += blockDim.x) { S.split[i] = 0; S.nmode[i] = 0; }
    for (int i = tid; i < 36; i += blockDim.x) S.p4idx[i] = k_P4IDX[i];
    if (tid == 0) S.quit = 0;
    __syncthreads();
    if (S.plane >= 4095) return;
    const fdvg_iplane P = planes[S.plane];
    const int uc = S.unit % P.cols, ur = S.unit / P.cols;
    UnitCtx U;
    U.x0 = uc * IU; U.y0 = ur * IU; U.w = P.w; U.h = P.h; U.qp = P.qp; U.lambda = P.lambda;
    const int vw = P.w - U.x0 < IU ? P.w - U.x0 : IU, vh = P.h - U.y0 < IU ? P.h - U.y0 : IU;
    // the unit's source, whole words (plane widths are multiples of 4)
    for (int i = tid; i < IU * IU / 4; i += blockDim.x) {
        const int yy = i / (IU / 4), xx = (i % (IU / 4)) * 4;
        uint32_t v = 0;
        if (yy < vh && xx < vw) v = *(const gmem_t<uint32_t> *)(P.src + (size_t)(U.y0 + yy) * P.w + U.x0 + xx);
        *(uint32_t *)&S.src[yy * IU + xx] = v;
    }
    for (int i = tid; i < (IU + 1) * CVS / 4; i += blockDim.x) ((uint32_t *)S.cv)[i] = 0;
    __syncthreads();
    // the neighbours' edges: lanes 0..15 the left unit's right column, 16..31
    // the bottom row above, 32 the above-left word with the corner, 33 the
    // above-right word with the top-right samples
    if (wv == 0) {
        const uint64_t tag = (uint64_t)step << 32;
        const uint64_t *eb = P.edge;
        const uint64_t *src_w = nullptr;
        if (L < 16 && uc > 0) src_w = eb + ((size_t)S.unit - 1) * 32 + 16 + L;
        else if (L >= 16 && L < 32 && ur > 0) src_w = eb + ((size_t)S.unit - P.cols) * 32 + (L - 16);
        else if (L == 32 && ur > 0 && uc > 0) src_w = eb + ((size_t)S.unit - P.cols - 1) * 32 + 15;
        else if (L == 33 && ur > 0 && uc + 1 < P.cols) src_w = eb + ((size_t)S.unit - P.cols + 1) * 32;
        uint64_t v = src_w ? __hip_atomic_load(src_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : tag;
        for (unsigned spins = 0; !__all((v >> 32) == (uint64_t)step); ++spins) {
            if (spins > (1u << 24)) { if (L == 0) atomicOr(ctl + 1, 1u); break; }
            __builtin_amdgcn_s_sleep(1);
            if (src_w && (v >> 32) != (uint64_t)step) v = __hip_atomic_load(src_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const uint32_t w4 = (uint32_t)v;
        if (src_w) {
            if (L < 16) {
#pragma unroll
                for (int k = 0; k < 4; ++k) S.cv[CV(4 * L + k, -1)] = (uint8_t)(w4 >> (8 * k));
            } else if (L < 32) {
                *(uint32_t *)&S.cv[CV(-1, 4 * (L - 16))] = w4;
            } else if (L == 32) {
                S.cv[CV(-1, -1)] = (uint8_t)(w4 >> 24);
            } else {
                *(uint32_t *)&S.cv[CV(-1, IU)] = w4;
            }
        }
    }
    __syncthreads();
    if (wv == 0) {
        walk<0>(S, U, 0, 0, ctl + 1);
        // publish the edges first: the bottom row, then the right column
        uint32_t w4 = 0;
        if (L < 16) w4 = *(const uint32_t *)&S.cv[CV(IU - 1, 4 * L)];
        else if (L < 32) {
            const int r0 = 4 * (L - 16);
#pragma unroll
            for (int k = 0; k < 4; ++k) w4 |= (uint32_t)S.cv[CV(r0 + k, IU - 1)] << (8 * k);
        }
        if (L < 32)
            __hip_atomic_store(P.edge + (size_t)S.unit * 32 + L, (uint64_t)step << 32 | w4, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (L == 0) __hip_atomic_store(&S.quit, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        coder_loop(S, U, wv - 1, ctl + 1);
    }
    __syncthreads();
    // the reconstruction, the tree and the mode-context map of the unit
    for (int i = tid; i < IU * IU / 4; i += blockDim.x) {
        const int yy = i / (IU / 4), xx = (i % (IU / 4)) * 4;
        if (yy < vh && xx < vw)
            *(gmem_t<uint32_t> *)(P.rec + (size_t)(U.y0 + yy) * P.w + U.x0 + xx) = *(const uint32_t *)&S.cv[CV(yy, xx)];
    }
    const fdvg_iplane *Pg = planes + S.plane;          // per-depth arrays indexed at run time
    for (int i = tid; i < 85; i += blockDim.x) {
        const int d = i < 1 ? 0 : i < 5 ? 1 : i < 21 ? 2 : 3, k = i - nbase(d), side = 1 << d;
        const int gx = uc * side + k % side, gy = ur * side + k / side;
        const int tw = Pg->tw[d], th = Pg->th[d], to = Pg->toff[d];
        if (gx < tw && gy < th) {
            const size_t o = (size_t)gy * tw + gx;
            P.tree[to + o] = S.split[i];
            P.tree[to + (size_t)tw * th + o] = S.nmode[i];
        }
    }
    for (int i = tid; i < 256; i += blockDim.x) {
        const int cx = uc * 16 + (i & 15), cy = ur * 16 + (i >> 4);
        if (cx < P.w / 4 && cy < P.h / 4) P.cmode[(size_t)cy * (P.w / 4) + cx] = S.cm[i];
    }
}

// The surviving leaves of a unit recoded from the final reconstruction into
// the level plane and the region flags.  One workgroup per unit (blockIdx.x
// indexes `order`, any order).
extern "C" __global__ void __launch_bounds__(64 * NWE)
k_intra_emit(const fdvg_iplane *__restrict__ planes, const uint32_t *__restrict__ order, int nunits) {
    __shared__ EmitLds S;
    const int tid = threadIdx.x, wv = tid >> 6, L = tid & 63;
    if ((int)blockIdx.x >= nunits) return;
    const uint32_t o = order[blockIdx.x];
    const int pidx = (int)(o >> 20), unit = (int)(o & 0xfffff);
    const fdvg_iplane P = planes[pidx];
    const fdvg_iplane *Pg = planes + pidx;             // per-depth arrays indexed at run time
    const int uc = unit % P.cols, ur = unit / P.cols, x0 = uc * IU, y0 = ur * IU;
    const int vw = P.w - x0 < IU ? P.w - x0 : IU, vh = P.h - y0 < IU ? P.h - y0 : IU;
    for (int i = tid; i < IU * IU / 4; i += blockDim.x) {
        const int yy = i / (IU / 4), xx = (i % (IU / 4)) * 4;
        uint32_t v = 0;
        if (yy < vh && xx < vw) v = *(const gmem_t<uint32_t> *)(P.src + (size_t)(y0 + yy) * P.w + x0 + xx);
        *(uint32_t *)&S.src[yy * IU + xx] = v;
    }
    // the final reconstruction with its borders: rows -1 .. 63, columns -1 .. 67
    for (int i = tid; i < (IU + 1) * (IU + 5); i += blockDim.x) {
        const int yy = i / (IU + 5) - 1, xx = i % (IU + 5) - 1;
        const int gx = x0 + xx, gy = y0 + yy;
        S.cv[CV(yy, xx)] = gx >= 0 && gy >= 0 && gx < P.w && gy < P.h ? P.rec[(size_t)gy * P.w + gx] : 0;
    }
    // the surviving leaves, by a walk of the decided tree (thread 0)
    if (tid == 0) {
        int nr = 0, nq = 0;
        int stk[32], sp = 0;                      // (depth << 12 | uy/4 << 6 | ux/4)
        stk[sp++] = 0;
        while (sp > 0) {
            const int e = stk[--sp], d = e >> 12, uy = ((e >> 6) & 63) * 4, ux = (e & 63) * 4;
            const int n = IU >> d;
            if (n == 4) { S.qjob[nq++] = (uint32_t)(uy << 8 | ux); continue; }
            const int gx = (x0 + ux) / n, gy = (y0 + uy) / n;
            const int tw = Pg->tw[d], th = Pg->th[d], to = Pg->toff[d];
            const size_t no = (size_t)gy * tw + gx;
            const int sp_ = P.tree[to + no];
            if (sp_ == 0) {
                const int m = P.tree[to + (size_t)tw * th + no];
                const int nreg = (n / 8) * (n / 8);
                for (int r = 0; r < nreg; ++r) S.rjob[nr++] = (uint32_t)(d << 28 | m << 24 | r << 16 | uy << 8 | ux);
                continue;
            }
            for (int k = 3; k >= 0; --k) {
                const int cx = ux + (k & 1) * (n / 2), cy = uy + (k >> 1) * (n / 2);
                if (x0 + cx < P.w && y0 + cy < P.h) stk[sp++] = (d + 1) << 12 | (cy / 4) << 6 | (cx / 4);
            }
        }
        S.nr = nr; S.nq = nq;
    }
    __syncthreads();
    const int nr = S.nr, nq = S.nq;
    // large-leaf regions: wave w takes jobs w, w + NWE, ...
    for (int j = wv; j < nr; j += NWE) {
        const uint32_t jb = S.rjob[j];
        const int d = jb >> 28, m = (jb >> 24) & 3, r = (jb >> 16) & 255, uy = (jb >> 8) & 255, ux = jb & 255;
        const int n = IU >> d, nrw = n / 8, rx = (r % nrw) * 8, ry = (r / nrw) * 8;
        int prm[8];
        leafn_setup(S.cv, ux, uy, n, x0 + ux, y0 + uy, S.top[wv], S.left[wv], prm, L);
        wave_sync();
        S.hpred[wv][L] = (uint8_t)pn(prm, S.top[wv], S.left[wv], m, rx + (L & 7), ry + (L >> 3));
        wave_sync();
        region_code<0>(S.ws, wv, &S.src[(uy + ry) * IU + ux + rx], IU, S.hpred[wv], 8, S.hrec[wv], 8,
                       &S.hlev[0][0], &S.hflag[0], P.qp, P.lambda);
        wave_sync();
        const int f = S.hflag[wv];
        // flag 1: four raster 4x4 blocks of 16; else one raster 8x8
        const int yy = f == 1 ? ((L >> 5) << 2) + ((L >> 2) & 3) : L >> 3;
        const int xx = f == 1 ? (((L >> 4) & 1) << 2) + (L & 3) : L & 7;
        P.lev[(size_t)(y0 + uy + ry + yy) * P.w + x0 + ux + rx + xx] = S.hlev[wv][L];
        if (L == 0) P.rflag[(size_t)((y0 + uy + ry) / 8) * ((P.w + 7) / 8) + (x0 + ux + rx) / 8] = (uint8_t)f;
        wave_sync();
    }
    // 4x4 leaves: their mode from the context map, the RDOQ coding again
    for (int j = wv; j < nq; j += NWE) {
        const int uy = S.qjob[j] >> 8, ux = S.qjob[j] & 255;
        const int bx = x0 + ux, by = y0 + uy;
        int ht, hl, htr, htl;
        const int dcv = edge4(S.cv, S.E[wv], ux, uy, bx, by, P.w, ht, hl, htr, htl, L);
        wave_sync();
        const int m = P.cmode[(size_t)(by / 4) * (P.w / 4) + bx / 4];
        uint32_t pw[4], sw[4];
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            uint32_t v = 0;
#pragma unroll
            for (int x = 0; x < 4; ++x) v |= (uint32_t)p4(S.E[wv], m, x, y, dcv) << (8 * x);
            pw[y] = v;
            sw[y] = *(const uint32_t *)&S.src[(uy + y) * IU + ux];
        }
        uint32_t rw[4];
        int lf[16], e, bits;
        final4(pw, sw, P.qp, P.lambda, S.rdj[wv], L, rw, lf, e, bits);
        if (L == 0) {
#pragma unroll
            for (int i = 0; i < 16; ++i) P.lev[(size_t)(by + (i >> 2)) * P.w + bx + (i & 3)] = (int16_t)lf[i];
        }
        wave_sync();
    }
}

extern "C" int fdvg_set_tables_intra(const int32_t *c8, const int32_t *qs8) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_T8), c8, 64 * sizeof(int32_t)) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_QS8), qs8, 52 * sizeof(int32_t)) != hipSuccess) return -1;
    return 0;
}

extern "C" int fdvg_launch_intra(const fdvg_iplane *planes, const uint32_t *order, int nunits, unsigned *ctl,
                                 unsigned step, hipStream_t st) {
    hipLaunchKernelGGL(k_intra_unit, dim3(nunits), dim3(64 * NWI), 0, st, planes, order, nunits, ctl, step);
    if (hipGetLastError() != hipSuccess) return -1;
    hipLaunchKernelGGL(k_intra_emit, dim3(nunits), dim3(64 * NWE), 0, st, planes, order, nunits);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------ host driver

#include <mutex>
#include <algorithm>
#include <vector>
#include "../../include/fdvgpu.h"

extern "C" void fdvg_seterr(const char *fmt, ...);
extern "C" void fdvg_tables8(int32_t c8[64], int32_t qs8[52]);
extern "C" int fdvg_launch_deblock(const void *planes, int nplanes, int maxpx, int bs, hipStream_t st);
extern "C" int fdvg_launch_zero(uint32_t *p, int n, hipStream_t st);

namespace {
struct IntraCtx {
    int dev = -1;
    hipStream_t st = nullptr;
    uint8_t *buf = nullptr;
    size_t cap = 0;
    unsigned step = 0;
};
IntraCtx g_ic;
std::mutex g_imu;

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
int cdiv(int a, int b) { return (a + b - 1) / b; }
}  // namespace

extern "C" size_t fdvgpu_itree_bytes(int w, int h) {
    if (w <= 0 || h <= 0) return 0;
    size_t t = 0;
    for (int d = 0; d < 4; ++d) t += 2 * (size_t)cdiv(w, IU >> d) * cdiv(h, IU >> d);
    return t;
}

#define ICK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    fdvg_seterr("%s: %s", #x, hipGetErrorString(e_)); return -1; } } while (0)

extern "C" int fdvgpu_image_batch(int n, const fdvgpu_iplane_in *in, fdvgpu_itree_out *out) {
    if (n < 1 || n > 4094 || !in || !out) { fdvg_seterr("fdvgpu_image_batch: bad arguments"); return -1; }
    size_t nunits = 0;
    int maxpx = 0;
    for (int i = 0; i < n; ++i) {
        const fdvgpu_iplane_in &p = in[i];
        if (!p.src || p.w <= 0 || p.h <= 0 || (p.w & 3) || (p.h & 3) || p.stride < p.w || p.qp < 0 || p.qp > 51 ||
            p.w > 65535 || p.h > 65535 || !out[i].rec || !out[i].lev || !out[i].rflag || !out[i].cmode ||
            !out[i].tree) {
            fdvg_seterr("fdvgpu_image_batch: plane %d: bad arguments", i);
            return -1;
        }
        nunits += (size_t)cdiv(p.w, IU) * cdiv(p.h, IU);
        maxpx = std::max(maxpx, p.w * p.h);
    }
    if (nunits >= (1u << 20)) { fdvg_seterr("fdvgpu_image_batch: too many units"); return -1; }
    std::lock_guard<std::mutex> lk(g_imu);
    IntraCtx &C = g_ic;
    int dev = 0;
    ICK(hipGetDevice(&dev));
    if (C.dev != dev) {
        if (C.buf) (void)hipFree(C.buf);
        C.buf = nullptr; C.cap = 0;
        if (C.st) (void)hipStreamDestroy(C.st);
        C.st = nullptr;
        ICK(hipStreamCreate(&C.st));
        int32_t c8[64], qs8[52];
        fdvg_tables8(c8, qs8);
        if (fdvg_set_tables_intra(c8, qs8) != 0) { fdvg_seterr("fdvgpu_image_batch: table upload failed"); return -1; }
        C.dev = dev;
    }
    // device layout: descriptors, order, ctl, deblock planes, the edge words
    // of every unit (cleared before each launch: a word counts only if it
    // carries this call's step, and the arena's other bytes are reused), then
    // per plane source, reconstruction, levels, flags, context map, tree
    std::vector<fdvg_iplane> P(n);
    std::vector<PlaneRef> R(n);
    const size_t edge_off = al256(n * sizeof(fdvg_iplane)) + al256(nunits * 4) + 256 + al256(n * sizeof(PlaneRef));
    size_t off = edge_off + al256(nunits * 32 * 8);
    std::vector<size_t> po(n);
    for (int i = 0; i < n; ++i) {
        const int w = in[i].w, h = in[i].h;
        const size_t px = (size_t)w * h;
        po[i] = off;
        off += al256(px) * 2 + al256(2 * px) + al256((size_t)cdiv(w, 8) * cdiv(h, 8)) + al256(px / 16) +
               al256(fdvgpu_itree_bytes(w, h));
    }
    if (off > C.cap) {
        if (C.buf) (void)hipFree(C.buf);
        C.buf = nullptr; C.cap = 0;
        const size_t want = off + off / 4;
        ICK(hipMalloc(&C.buf, want));
        C.cap = want;
    }
    uint8_t *B = C.buf;
    fdvg_iplane *P_d = (fdvg_iplane *)B;
    uint32_t *order_d = (uint32_t *)(B + al256(n * sizeof(fdvg_iplane)));
    unsigned *ctl_d = (unsigned *)((uint8_t *)order_d + al256(nunits * 4));
    PlaneRef *R_d = (PlaneRef *)((uint8_t *)ctl_d + 256);
    uint64_t *edge_d = (uint64_t *)(B + edge_off);
    size_t ubase = 0;
    if (++C.step == 0) C.step = 1;
    std::vector<uint64_t> keys;
    keys.reserve(nunits);
    for (int i = 0; i < n; ++i) {
        const int w = in[i].w, h = in[i].h;
        const size_t px = (size_t)w * h;
        uint8_t *q = B + po[i];
        fdvg_iplane &D = P[i];
        D.src = q; q += al256(px);
        D.rec = q; q += al256(px);
        D.lev = (int16_t *)q; q += al256(2 * px);
        D.rflag = q; q += al256((size_t)cdiv(w, 8) * cdiv(h, 8));
        D.cmode = q; q += al256(px / 16);
        D.tree = q;
        D.edge = edge_d + ubase * 32;
        D.lambda = 0.425 * pow(2.0, (in[i].qp - 12) / 3.0);              // FDV_LAMBDA0 (fdv.h:3372), 4883
        D.w = w; D.h = h; D.qp = in[i].qp; D.cols = cdiv(w, IU); D.rows = cdiv(h, IU); D.pad = 0;
        int to = 0;
        for (int d = 0; d < 4; ++d) {
            D.toff[d] = to; D.tw[d] = cdiv(w, IU >> d); D.th[d] = cdiv(h, IU >> d);
            to += 2 * D.tw[d] * D.th[d];
        }
        R[i] = PlaneRef{D.rec, w, h, w, in[i].qp};
        ubase += (size_t)D.cols * D.rows;
        for (int u = 0; u < D.cols * D.rows; ++u) {
            const uint64_t key = (uint64_t)(u % D.cols + 2 * (u / D.cols));
            keys.push_back(key << 40 | (uint64_t)i << 20 | (uint64_t)u);
        }
        ICK(hipMemcpy2DAsync((void *)D.src, w, in[i].src, in[i].stride, w, h, hipMemcpyHostToDevice, C.st));
    }
    std::sort(keys.begin(), keys.end());
    std::vector<uint32_t> order(nunits);
    for (size_t k = 0; k < nunits; ++k) order[k] = (uint32_t)(keys[k] & 0xffffffffffull);
    ICK(hipMemcpyAsync(P_d, P.data(), n * sizeof(fdvg_iplane), hipMemcpyHostToDevice, C.st));
    ICK(hipMemcpyAsync(order_d, order.data(), nunits * 4, hipMemcpyHostToDevice, C.st));
    ICK(hipMemcpyAsync(R_d, R.data(), n * sizeof(PlaneRef), hipMemcpyHostToDevice, C.st));
    if (fdvg_launch_zero(ctl_d, 8, C.st) != 0 || fdvg_launch_zero((uint32_t *)edge_d, (int)(nunits * 64), C.st) != 0 ||
        fdvg_launch_intra(P_d, order_d, (int)nunits, ctl_d, C.step, C.st) != 0 ||
        fdvg_launch_deblock(R_d, n, maxpx, 256, C.st) != 0) {
        fdvg_seterr("fdvgpu_image_batch: launch failed");
        return -1;
    }
    unsigned ctl_h[2] = {0, 0};
    for (int i = 0; i < n; ++i) {
        const int w = in[i].w, h = in[i].h;
        const size_t px = (size_t)w * h;
        ICK(hipMemcpyAsync(out[i].rec, P[i].rec, px, hipMemcpyDeviceToHost, C.st));
        ICK(hipMemcpyAsync(out[i].lev, P[i].lev, 2 * px, hipMemcpyDeviceToHost, C.st));
        ICK(hipMemcpyAsync(out[i].rflag, P[i].rflag, (size_t)cdiv(w, 8) * cdiv(h, 8), hipMemcpyDeviceToHost, C.st));
        ICK(hipMemcpyAsync(out[i].cmode, P[i].cmode, px / 16, hipMemcpyDeviceToHost, C.st));
        ICK(hipMemcpyAsync(out[i].tree, P[i].tree, fdvgpu_itree_bytes(w, h), hipMemcpyDeviceToHost, C.st));
    }
    ICK(hipMemcpyAsync(ctl_h, ctl_d, sizeof ctl_h, hipMemcpyDeviceToHost, C.st));
    ICK(hipStreamSynchronize(C.st));
    if (ctl_h[1]) { fdvg_seterr("fdvgpu_image_batch: a bounded wait ran out (flags %u)", ctl_h[1]); return -1; }
    return 0;
}

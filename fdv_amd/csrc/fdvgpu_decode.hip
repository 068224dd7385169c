// fdvgpu_decode.hip -- P-frame reconstruction from parsed symbols on gfx950:
// the device half of the reference's pframe_decode (yhbsh/fdv include/fdv.h
// 6492-6737), given each MB's decision and quantised levels (the host parses
// the entropy-coded payload into them).
//
// Every non-intra MB depends on nothing but its own record and the reference
// frames, so k_dinter reconstructs all of them at once (one workgroup per MB,
// XCD column strips).  Intra MBs read their left, top and top-left neighbours'
// unfiltered reconstruction: k_dintra takes them in anti-diagonal order (the
// list k_plist builds) on persistent workgroups, reading inter neighbours from
// the output planes and intra ones from their handoff words.  The band deblock
// then runs on the planes (k_deblock, as after encoding).
//
// Prediction samples come from small tiles of the reference staged in LDS at
// the vector's integer position, so any vector the reference decoder accepts
// (mv_in_bounds, fdv.h:6480: within the planes' 32-sample border) is read in
// bounds; the host parser rejects the others, as the reference does.
#include <hip/hip_runtime.h>
#include "fdvgpu_common.h"
#include "fdvgpu_dev.h"
#include "fdvgpu_resid.h"
#include "fdvgpu_mb.h"

#define DT_S 24                 // luma tile stride: a 16x16 block needs 21 x 21 samples
#define DT_N 21

// Chroma tiles: 81 bytes for a 9x9, 4 x 25 for the INTER8 quadrants' 5x5s.
static_assert(4 * 25 <= 112 && 81 <= 112, "chroma tiles");
struct __align__(16) DSmem {
    fdvg_mbrec R;
    uint8_t tile[4][DT_S * DT_N];        // SKIP / INTER16: tile 0; INTER8: one per quadrant
    uint8_t ctile[2][112];               // per plane: one 9x9 tile (stride 9), or four 5x5 (one per quadrant)
    uint8_t pred[256], cpred[128];
    uint8_t rec[256], crec[128];
    uint8_t hbytes[17], lcol[32], ctop[2][9];
    int32_t x32[6][64];
    int64_t x64[6][64];
};

// H.264 quarter-pel luma sample (fdv.h:2931-2990, fig. 8-4) from a tile of
// stride S whose integer sample for this output pixel is at (X, Y).
template <int S> __device__ __forceinline__ int d_hr(const uint8_t *w, int x, int y) {
    const uint8_t *p = w + y * S + x;
    return tap6(p[-2], p[-1], p[0], p[1], p[2], p[3]);
}
template <int S> __device__ __forceinline__ int d_vr(const uint8_t *w, int x, int y) {
    const uint8_t *p = w + y * S + x;
    return tap6(p[-2 * S], p[-S], p[0], p[S], p[2 * S], p[3 * S]);
}
template <int S> __device__ __forceinline__ int d_b(const uint8_t *w, int x, int y) { return clip8((d_hr<S>(w, x, y) + 16) >> 5); }
template <int S> __device__ __forceinline__ int d_h(const uint8_t *w, int x, int y) { return clip8((d_vr<S>(w, x, y) + 16) >> 5); }
template <int S> __device__ __forceinline__ int d_j(const uint8_t *w, int x, int y) {
    const int v = tap6(d_hr<S>(w, x, y - 2), d_hr<S>(w, x, y - 1), d_hr<S>(w, x, y), d_hr<S>(w, x, y + 1),
                       d_hr<S>(w, x, y + 2), d_hr<S>(w, x, y + 3));
    return clip8((v + 512) >> 10);
}
template <int S> __device__ __forceinline__ int tile_luma(const uint8_t *w, int X, int Y, int ph) {
    const int G = w[Y * S + X];
    switch (ph) {
    case 0:  return G;
    case 1:  return av2(G, d_b<S>(w, X, Y));
    case 2:  return d_b<S>(w, X, Y);
    case 3:  return av2(d_b<S>(w, X, Y), w[Y * S + X + 1]);
    case 4:  return av2(G, d_h<S>(w, X, Y));
    case 5:  return av2(d_b<S>(w, X, Y), d_h<S>(w, X, Y));
    case 6:  return av2(d_b<S>(w, X, Y), d_j<S>(w, X, Y));
    case 7:  return av2(d_b<S>(w, X, Y), d_h<S>(w, X + 1, Y));
    case 8:  return d_h<S>(w, X, Y);
    case 9:  return av2(d_h<S>(w, X, Y), d_j<S>(w, X, Y));
    case 10: return d_j<S>(w, X, Y);
    case 11: return av2(d_j<S>(w, X, Y), d_h<S>(w, X + 1, Y));
    case 12: return av2(d_h<S>(w, X, Y), w[(Y + 1) * S + X]);
    case 13: return av2(d_h<S>(w, X, Y), d_b<S>(w, X, Y + 1));
    case 14: return av2(d_j<S>(w, X, Y), d_b<S>(w, X, Y + 1));
    default: return av2(d_h<S>(w, X + 1, Y), d_b<S>(w, X, Y + 1));
    }
}

// Stage an n x n tile (row stride `ts`) of plane p (stride ps) whose (0, 0) is
// the plane's sample (x0, y0); threads [t0, t0 + nthr) share the work.
__device__ __forceinline__ void stage_tile(uint8_t *dst, int ts, const uint8_t *p, int ps, int x0, int y0, int n,
                                           int t, int nthr) {
    const gmem_t<uint8_t> *g = gmem(p);
    for (int k = t; k < n * n; k += nthr) dst[(k / n) * ts + k % n] = g[(ptrdiff_t)(y0 + k / n) * ps + x0 + k % n];
}

// The residual of one 4x4 block: dequantise and inverse-transform levels
// (raster order) onto the prediction (fdv.h:2546, 2275).
__device__ __forceinline__ void add_block4(const int16_t *lev, const uint8_t *pred, int pst, uint8_t *rec, int rst,
                                           int qp) {
    const int m6 = qp % 6, sh = 1 << (qp / 6);
    const int dv0 = k_DV[m6][0], dv1 = k_DV[m6][1], dv2 = k_DV[m6][2];
    int dq[16], rr[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int dv = cls4(i) == 0 ? dv0 : cls4(i) == 1 ? dv1 : dv2;
        dq[i] = ((int)lev[i] * dv) * sh;
    }
    idct4_reg(dq, rr);
#pragma unroll
    for (int i = 0; i < 16; ++i) rec[(i >> 2) * rst + (i & 3)] = (uint8_t)clip8(pred[(i >> 2) * pst + (i & 3)] + rr[i]);
}

// The residual of one 8x8 block (dequant8 saturating to int32, int64 inverse
// with 2^35 rounding, fdv.h:2639-2680) by the 64 lanes of one wavefront.
__device__ __forceinline__ void add_block8(DSmem &s, int slot, const int16_t *lev, const uint8_t *pred, int pst,
                                           uint8_t *rec, int rst, int qp) {
    const int L = threadIdx.x & 63, y8 = L >> 3, x8 = L & 7;
    const long long dqv = (long long)lev[L] * c_QS8[qp];
    s.x32[slot][L] = dqv > 2147483647LL ? 2147483647 : dqv < -2147483648LL ? (-2147483647 - 1) : (int)dqv;
    wave_sync();
    long long B = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) B += (long long)c_T8[k * 8 + y8] * s.x32[slot][k * 8 + x8];
    s.x64[slot][L] = B;
    wave_sync();
    long long X = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) X += s.x64[slot][y8 * 8 + k] * (long long)c_T8[k * 8 + x8];
    const int rr = (int16_t)((X + (1LL << 35)) >> 36);
    rec[y8 * rst + x8] = (uint8_t)clip8(pred[y8 * pst + x8] + rr);
}

// Luma prediction and chroma prediction of a non-intra MB into s.pred / s.cpred.
__device__ __forceinline__ void dec_predict_inter(DSmem &s, const fdvg_seq &q, int bx, int by) {
    const int tid = threadIdx.x, cbx = bx >> 1, cby = by >> 1, mode = s.R.mode;
    if (mode != 3) {
        const int r = s.R.ref, mx = s.R.mv[0][0], my = s.R.mv[0][1];
        stage_tile(s.tile[0], DT_S, r ? q.ry[1] : q.ry[0], q.ys, bx + (mx >> 2) - 2, by + (my >> 2) - 2, DT_N, tid,
                   256);
        // chroma vector: the luma vector / 2, truncated toward zero (fdv.h:6061)
        const int cmx = mx / 2, cmy = my / 2;
        if (tid < 162) {
            const int pl = tid >= 81, k = tid - 81 * pl;
            const gmem_t<uint8_t> *p = gmem(pl ? (r ? q.rv[1] : q.rv[0]) : (r ? q.ru[1] : q.ru[0]));
            s.ctile[pl][k] = p[(ptrdiff_t)(cby + (cmy >> 2) + k / 9) * q.cs + cbx + (cmx >> 2) + k % 9];
        }
        __syncthreads();
        s.pred[tid] = (uint8_t)tile_luma<DT_S>(s.tile[0], (tid & 15) + 2, (tid >> 4) + 2, ((my & 3) << 2) | (mx & 3));
        if (tid < 128) {
            const int pl = tid >> 6, x = tid & 7, y = (tid >> 3) & 7, fx = cmx & 3, fy = cmy & 3;
            const uint8_t *t = &s.ctile[pl][y * 9 + x];
            const int top = t[0] * (4 - fx) + t[1] * fx, bot = t[9] * (4 - fx) + t[10] * fx;
            s.cpred[tid] = (uint8_t)((top * (4 - fy) + bot * fy + 8) >> 4);
        }
    } else {
        // INTER8: each quadrant from its own vector on the nearest reference
        for (int qd = 0; qd < 4; ++qd) {
            const int mx = s.R.mv[qd][0], my = s.R.mv[qd][1], qx = (qd & 1) * 8, qy = (qd >> 1) * 8;
            stage_tile(s.tile[qd], 16, q.ry[0], q.ys, bx + qx + (mx >> 2) - 2, by + qy + (my >> 2) - 2, 13, tid, 256);
            const int cmx = mx / 2, cmy = my / 2;
            if (tid < 50) {
                const int pl = tid >= 25, k = tid - 25 * pl;
                const gmem_t<uint8_t> *p = gmem(pl ? q.rv[0] : q.ru[0]);
                s.ctile[pl][25 * qd + k] = p[(ptrdiff_t)(cby + qy / 2 + (cmy >> 2) + k / 5) * q.cs + cbx + qx / 2 + (cmx >> 2) + k % 5];
            }
        }
        __syncthreads();
        {
            const int x = tid & 15, y = tid >> 4, qd = (y >> 3) * 2 + (x >> 3);
            const int mx = s.R.mv[qd][0], my = s.R.mv[qd][1];
            s.pred[tid] = (uint8_t)tile_luma<16>(s.tile[qd], (x & 7) + 2, (y & 7) + 2, ((my & 3) << 2) | (mx & 3));
        }
        if (tid < 128) {
            const int pl = tid >> 6, x = tid & 7, y = (tid >> 3) & 7, qd = (y >> 2) * 2 + (x >> 2);
            const int cmx = s.R.mv[qd][0] / 2, cmy = s.R.mv[qd][1] / 2, fx = cmx & 3, fy = cmy & 3;
            const uint8_t *t = &s.ctile[pl][25 * qd + (y & 3) * 5 + (x & 3)];
            const int top = t[0] * (4 - fx) + t[1] * fx, bot = t[5] * (4 - fx) + t[6] * fx;
            s.cpred[tid] = (uint8_t)((top * (4 - fy) + bot * fy + 8) >> 4);
        }
    }
    __syncthreads();
}

// The residual onto the prediction (decode_residual, fdv.h:5426; chroma per
// the MB's chroma kind, 6429-6470), into s.rec / s.crec.
__device__ __forceinline__ void dec_residual(DSmem &s, int qp) {
    const int tid = threadIdx.x, wv = tid >> 6, L = tid & 63;
    const fdvg_mbrec &R = s.R;
    // luma: region wv (8x8 at raster position wv) by wavefront wv
    {
        const int o = (wv >> 1) * 128 + (wv & 1) * 8;
        const int fl = R.mode == 0 ? 0 : R.lflag[wv];
        if (fl == 2) add_block8(s, wv, R.lev + 64 * wv, s.pred + o, 16, s.rec + o, 16, qp);
        else {
            if (L < 4 && fl == 1) {
                const int bo = o + (L >> 1) * 64 + (L & 1) * 4;
                add_block4(R.lev + 64 * wv + 16 * L, s.pred + bo, 16, s.rec + bo, 16, qp);
            }
            if (fl == 0) { const int y = L >> 3, x = L & 7; s.rec[o + y * 16 + x] = s.pred[o + y * 16 + x]; }
        }
    }
    // chroma: plane p by wavefront p (waves 0, 1 after their luma region)
    __syncthreads();
    if (wv < 2) {
        const int p = wv, ck = R.ckind;
        const int fl = ck == 0 ? 0 : ck == 3 ? 1 : R.cflag[p];       // INTER16 any-flag: 0 / 1
        const uint8_t *cp = s.cpred + 64 * p;
        uint8_t *cr = s.crec + 64 * p;
        if (fl == 2) add_block8(s, 4 + p, R.clev + 64 * p, cp, 8, cr, 8, qp);
        else {
            if (L < 4 && fl == 1) {
                const int bo = (L >> 1) * 32 + (L & 1) * 4;
                add_block4(R.clev + 64 * p + 16 * L, cp + bo, 8, cr + bo, 8, qp);
            }
            if (fl == 0) cr[L] = cp[L];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void dec_write(const DSmem &s, const fdvg_seq &q, int bx, int by) {
    const int tid = threadIdx.x, cbx = bx >> 1, cby = by >> 1;
    gmemw(q.oy)[(ptrdiff_t)(by + (tid >> 4)) * q.ys + bx + (tid & 15)] = s.rec[tid];
    if (tid < 128) {
        const int pl = tid >> 6, p = tid & 63;
        gmemw(pl ? q.ov : q.ou)[(ptrdiff_t)(cby + (p >> 3)) * q.cs + cbx + (p & 7)] = s.crec[tid];
    }
}

__device__ __forceinline__ void dec_load_rec(DSmem &s, const fdvg_mbrec *R) {
    static_assert(sizeof(fdvg_mbrec) % 16 == 0 && sizeof(fdvg_mbrec) / 16 <= 256, "record staging");
    if (threadIdx.x < (int)(sizeof(fdvg_mbrec) / 16)) ((uint4 *)&s.R)[threadIdx.x] = ((const uint4 *)R)[threadIdx.x];
    __syncthreads();
}

extern "C" __global__ void __launch_bounds__(256)
k_dinter(const fdvg_seq *__restrict__ seqs, const fdvg_mbrec *__restrict__ recs) {
    __shared__ DSmem s;
    const fdvg_seq q = seqs[blockIdx.y];
    int i;
    if (!xcd_mb(blockIdx.x, q.w >> 4, q.h >> 4, i)) return;
    const int mb = q.mb0 + i, mw = q.w >> 4, bx = (i % mw) * 16, by = (i / mw) * 16;
    if (recs[mb].mode == 2) return;
    dec_load_rec(s, &recs[mb]);
    dec_predict_inter(s, q, bx, by);
    dec_residual(s, q.qp);
    dec_write(s, q, bx, by);
}

// An intra MB's neighbours (gather_nb, fdv.h:5201-5212): luma corner + top row
// in s.hbytes, left column in s.lcol[0..15], chroma corners + top rows in
// s.ctop, left columns in s.lcol[16..31]; intra neighbours from their handoff
// words (the frame step's id `tag`), the others from the output planes.
__device__ __forceinline__ void dec_neighbours(DSmem &s, const fdvg_seq &q, const fdvg_mbrec *recs,
                                               const uint64_t *ih, unsigned tag, unsigned *ctl, int mx, int my,
                                               int mw, int bx, int by) {
    const int tid = threadIdx.x, cbx = bx >> 1, cby = by >> 1, ht = my > 0, hl = mx > 0;
    const int i_top = (my - 1) * mw + mx, i_left = my * mw + mx - 1, i_tl = (my - 1) * mw + mx - 1;
    const fdvg_mbrec *R0 = recs + q.mb0;
    const bool top_i = ht && R0[i_top].mode == 2, left_i = hl && R0[i_left].mode == 2;
    const bool tl_i = ht && hl && R0[i_tl].mode == 2;
    const uint64_t *ih0 = ih + (size_t)q.mb0 * IH_WORDS;
    const gmem_t<uint8_t> *py = gmem(q.oy), *pu = gmem(q.ou), *pv = gmem(q.ov);
    const int ys = q.ys, cs = q.cs;
    if (tid < 16) {
        s.hbytes[1 + tid] = !ht ? 0 : top_i ? ih_byte(ih0 + (size_t)i_top * IH_WORDS, tid >> 2, tid & 3, tag, ctl)
                                           : py[(ptrdiff_t)(by - 1) * ys + bx + tid];
    } else if (tid == 16) {
        s.hbytes[0] = !(ht && hl) ? 128 : tl_i ? ih_byte(ih0 + (size_t)i_tl * IH_WORDS, 3, 3, tag, ctl)
                                               : py[(ptrdiff_t)(by - 1) * ys + bx - 1];
    } else if (tid >= 32 && tid < 48) {
        const int k = tid - 32;
        s.lcol[k] = !hl ? 0 : left_i ? ih_byte(ih0 + (size_t)i_left * IH_WORDS, 8 + (k >> 2), k & 3, tag, ctl)
                                     : py[(ptrdiff_t)(by + k) * ys + bx - 1];
    } else if (tid >= 64 && tid < 82) {
        const int pl = (tid - 64) / 9, k = (tid - 64) % 9;
        const gmem_t<uint8_t> *p = pl ? pv : pu;
        if (k == 0)
            s.ctop[pl][0] = !(ht && hl) ? 128 : tl_i ? ih_byte(ih0 + (size_t)i_tl * IH_WORDS, pl ? 7 : 5, 3, tag, ctl)
                                                     : p[(ptrdiff_t)(cby - 1) * cs + cbx - 1];
        else
            s.ctop[pl][k] = !ht ? 0 : top_i ? ih_byte(ih0 + (size_t)i_top * IH_WORDS, (pl ? 6 : 4) + ((k - 1) >> 2),
                                                      (k - 1) & 3, tag, ctl)
                                            : p[(ptrdiff_t)(cby - 1) * cs + cbx + k - 1];
    } else if (tid >= 96 && tid < 112) {
        const int pl = (tid - 96) >> 3, k = (tid - 96) & 7;
        const gmem_t<uint8_t> *p = pl ? pv : pu;
        s.lcol[16 + 8 * pl + k] = !hl ? 0 : left_i ? ih_byte(ih0 + (size_t)i_left * IH_WORDS, (pl ? 14 : 12) + (k >> 2),
                                                             k & 3, tag, ctl)
                                                   : p[(ptrdiff_t)(cby + k) * cs + cbx - 1];
    }
    __syncthreads();
}

extern "C" __global__ void __launch_bounds__(256)
k_dintra(const fdvg_seq *__restrict__ seqs, int nseq, const fdvg_mbrec *__restrict__ recs, unsigned *ctl,
         const unsigned *__restrict__ ilist, uint64_t *__restrict__ ih) {
    __shared__ DSmem s;
    __shared__ int job_s;
    const int tid = threadIdx.x;
    const unsigned n = ilist[0];
    for (;;) {
        if (tid == 0) job_s = (int)atomicAdd(ctl + 5, 1u);
        __syncthreads();
        const unsigned job = (unsigned)__builtin_amdgcn_readfirstlane(job_s);   // uniform: scalar registers
        __syncthreads();
        if (job >= n) return;
        const int mb = __builtin_amdgcn_readfirstlane((int)ilist[1 + job]);
        int sq = 0;
        for (int k = 1; k < nseq; ++k) if (mb >= seqs[k].mb0) sq = k;
        const fdvg_seq q = seqs[sq];
        const int i = mb - q.mb0, mw = q.w >> 4, mx = i % mw, my = i / mw, bx = mx * 16, by = my * 16;
        dec_load_rec(s, &recs[mb]);
        dec_neighbours(s, q, recs, ih, q.step, ctl, mx, my, mw, bx, by);
        // 16x16 luma and 8x8 chroma from the MB's sub-mode (fdv.h:6536-6552)
        const int sub = s.R.sub;
        {
            const Nb nb = nb_setup(&s.hbytes[1], s.lcol, s.hbytes[0], 16, my > 0, mx > 0);
            s.pred[tid] = (uint8_t)ipred(sub, nb, &s.hbytes[1], s.lcol, 16, tid & 15, tid >> 4);
        }
        if (tid < 128) {
            const int pl = tid >> 6, p = tid & 63;
            const uint8_t *tp = &s.ctop[pl][1], *lf = &s.lcol[16 + 8 * pl];
            const Nb nb = nb_setup(tp, lf, s.ctop[pl][0], 8, my > 0, mx > 0);
            s.cpred[tid] = (uint8_t)ipred(sub, nb, tp, lf, 8, p & 7, p >> 3);
        }
        __syncthreads();
        dec_residual(s, q.qp);
        dec_write(s, q, bx, by);
        ih_publish(ih + (size_t)mb * IH_WORDS, s.rec, s.crec, q.step);
        __syncthreads();
    }
}

// This file's copy of the 8x8 tables (constant memory is per code object).
extern "C" int fdvg_set_tables_decode(const int32_t *c8, const int32_t *qs8) {
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_T8), c8, 64 * sizeof(int32_t)) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_QS8), qs8, 52 * sizeof(int32_t)) != hipSuccess) return -1;
    return 0;
}

extern "C" int fdvg_launch_plist(const fdvg_seq *seqs, int nseq, const fdvg_mbinfo *given, int total_mb,
                                 unsigned *ilist, hipStream_t st);

// One P-frame of every sequence from its MB records: the intra list, the
// inter MBs, then the intra MBs (job counter ctl[5], zeroed by the caller).
extern "C" int fdvg_launch_decode(const fdvg_seq *seqs, int nseq, int xgrid, const fdvg_mbinfo *modes, int total_mb,
                                  const fdvg_mbrec *recs, unsigned *ctl, unsigned *ilist, uint64_t *ih,
                                  hipStream_t st) {
    static int cap_i = 0;
    if (!cap_i) {
        int dev = 0, ncu = 0, pi = 0;
        if (hipGetDevice(&dev) != hipSuccess) return -1;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pi, (const void *)k_dintra, 256, 0) != hipSuccess) return -1;
        cap_i = ncu * (pi > 0 ? pi : 1);
    }
    if (fdvg_launch_plist(seqs, nseq, modes, total_mb, ilist, st) != 0) return -1;
    hipLaunchKernelGGL(k_dinter, dim3(xgrid, nseq), dim3(256), 0, st, seqs, recs);
    hipLaunchKernelGGL(k_dintra, dim3(total_mb < cap_i ? total_mb : cap_i), dim3(256), 0, st, seqs, nseq, recs, ctl,
                       (const unsigned *)ilist, ih);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// fdvgpu_metab.hip -- block-parallel motion tables for the exhaustive search mode.
//
// The reference's EPZS diamond (fdv.h:3375-3466) starts from the running vector
// predictor, so it can only run inside the MB wavefront.  The exhaustive search
// scores every whole-pel vector within +-16 pels (1089 points) and so depends on
// nothing but the source block and the reference: it runs for all MBs of all
// sequences at once, before the wavefront, and the wavefront only reads its
// results.  Ties go to the smaller |dx|+|dy|, then to raster order.  The best
// point is then refined with the reference's own sub-pel step (half- then
// quarter-pel four-point cross, unclamped, skipped at SAD 0: fdv.h:3453-3461).
//
// One 256-thread workgroup per MB; the source MB and a 64x64 window of each
// reference are staged in LDS.  Per reference-0 point the four 8x8 quadrant
// SADs are accumulated separately (the INTER8 searches) and summed for 16x16.
// The full search scores four horizontal offsets at once with
// v_qsad_pk_u16_u8 on aligned window words; the candidate search uses v_sad_u8
// on packed bytes, with v_alignbyte for unaligned windows.
#include <hip/hip_runtime.h>
#include "fdvgpu_common.h"
#include "fdvgpu_dev.h"

struct __align__(16) MSmem {
    uint8_t win[2][WIN * WIN];
    uint32_t cur32[64];                 // the source MB, 4 words a row
    uint8_t cur[256];
    unsigned long long key[6][4];       // per block, per wave: best packed key
    int32_t sad[2][6][4];               // per step parity, block, cross point
    int32_t res[6][3];                  // per block: mvx, mvy, sad after sub-pel
};

__device__ __forceinline__ uint32_t win_word(const uint8_t *w, int y, int x) {
    // 4 bytes at an arbitrary x of row y, from two aligned words
    const uint32_t *r = (const uint32_t *)(w + y * WIN);
    const int a = x >> 2, sh = x & 3;
    const uint32_t lo = r[a], hi = r[a + 1];
    return __builtin_amdgcn_alignbyte(hi, lo, sh);
}

__device__ __forceinline__ unsigned long long umin64(unsigned long long a, unsigned long long b) { return a < b ? a : b; }

// SAD of 4 predicted samples against the source MB: block-relative column x,
// rows y0 .. y0 + 3 of the block at (ox, oy) in the MB, window w, quarter-pel
// vector (px, py).
__device__ __forceinline__ int sad4_q(const uint8_t *w, const uint8_t *cur, int ox, int oy, int x, int y0, int px,
                                      int py) {
    return qerr4<WIN, false>(w, WOFF + ox + x + (px >> 2), WOFF + oy + y0 + (py >> 2), ((py & 3) << 2) | (px & 3),
                             cur + (oy + y0) * 16 + ox + x, 16);
}

// The reference's sub-pel step (half- then quarter-pel four-point cross,
// unclamped, skipped once a block's SAD is 0: fdv.h:3453-3461) for all six
// blocks at once (0, 1: 16x16 on references 0 and 1; 2..5: the 8x8 quadrants
// on reference 0): at each step wave k scores cross point k of every block
// still refining (16x16: a lane per column and 4 rows; quadrants: all four in
// one pass, 16 lanes each), then every block takes its best point (strict <,
// in cross order).  Each block's result is that of refining it alone.
__device__ __forceinline__ void refine_all(MSmem &s, int nr, int *bmx, int *bmy, int *bbest) {
    const int t = threadIdx.x, k = t >> 6, tt = t & 63;
    int par = 0;
#pragma unroll
    for (int st = 2; st >= 1; st >>= 1) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            if (bbest[b] <= 0 || (b == 1 && nr < 2)) continue;        // uniform
            const int px = bmx[b] + k_CROSS[k][0] * st, py = bmy[b] + k_CROSS[k][1] * st;
            int acc = sad4_q(s.win[b], s.cur, 0, 0, tt & 15, (tt >> 4) * 4, px, py);
            acc = gsum<64>(acc);
            if (tt == 0) s.sad[par][b][k] = acc;
        }
        if (bbest[2] > 0 || bbest[3] > 0 || bbest[4] > 0 || bbest[5] > 0) {   // uniform
            const int qd = tt >> 4;
            const int mx = qd == 0 ? bmx[2] : qd == 1 ? bmx[3] : qd == 2 ? bmx[4] : bmx[5];
            const int my = qd == 0 ? bmy[2] : qd == 1 ? bmy[3] : qd == 2 ? bmy[4] : bmy[5];
            int acc = sad4_q(s.win[0], s.cur, (qd & 1) * 8, (qd >> 1) * 8, tt & 7, ((tt >> 3) & 1) * 4,
                             mx + k_CROSS[k][0] * st, my + k_CROSS[k][1] * st);
            acc = gsum<16>(acc);
            if ((tt & 15) == 0) s.sad[par][2 + qd][k] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            if (bbest[b] <= 0 || (b == 1 && nr < 2)) continue;
            const int cx = bmx[b], cy = bmy[b];
            for (int k2 = 0; k2 < 4; ++k2) {
                const int v = s.sad[par][b][k2];
                if (v < bbest[b]) { bbest[b] = v; bmx[b] = cx + k_CROSS[k2][0] * st; bmy[b] = cy + k_CROSS[k2][1] * st; }
            }
        }
        par ^= 1;
    }
}

// With a proposer output `nn` whose confidence is at least `thr`, the integer
// search covers only the 3x3 whole-pel neighbourhoods of its vectors -- on the
// near reference the network's 8, the correlation map's 4 and the quadrant
// maps' 8, on the far one the first 12 doubled -- and the zero vector, same tie
// order; below `thr` it is the full search.  A SKIP probability of at least
// `sthr` sets flags bit 0.
extern "C" __global__ void __launch_bounds__(256)
k_me_exhaustive(const fdvg_seq *__restrict__ seqs, fdvg_metab *__restrict__ out,
                const fdvg_nnout *__restrict__ nn, float thr, float sthr) {
    __shared__ MSmem s;
    const fdvg_seq q = seqs[blockIdx.y];
    const int mw = q.w >> 4, nmb = mw * (q.h >> 4);
    int i;
    if (!xcd_mb(blockIdx.x, mw, q.h >> 4, i)) return;
    (void)nmb;
    const int tid = threadIdx.x, bx = (i % mw) * 16, by = (i / mw) * 16;
    const int nr = q.navail;
    s.cur[tid] = q.cy[(size_t)(by + (tid >> 4)) * q.w + bx + (tid & 15)];
    for (int r = 0; r < nr; ++r) {
        const int wy = tid >> 2, wx = (tid & 3) * 16;
        const uint8_t *src = (r ? q.ry[1] : q.ry[0]) + (ptrdiff_t)(by - WOFF + wy) * q.ys + bx - WOFF + wx;
        *(uint64_t *)&s.win[r][wy * WIN + wx] = *(const uint64_t *)src;
        *(uint64_t *)&s.win[r][wy * WIN + wx + 8] = *(const uint64_t *)(src + 8);
    }
    __syncthreads();
    if (tid < 64) s.cur32[tid] = ((const uint32_t *)s.cur)[tid];
    __syncthreads();

    // integer search: thread t takes points t, t+256, ...; packed key
    // (sad, |dx|+|dy|, raster index) so the minimum is the reference tie order.
    unsigned long long best[6];
    for (int b = 0; b < 6; ++b) best[b] = ~0ull;
    const fdvg_nnout *pr = nn ? &nn[q.mb0 + i] : nullptr;
    const bool cand = pr && pr->conf >= thr;
    // candidate mode: 98 rows of 3x3 whole-pel neighbourhoods, two threads a
    // row (source rows 0-7 | 8-15).  Row j < 60 is row j % 3 of the
    // neighbourhood of near-reference vector j / 3 (the network's 8, the map's
    // 4, the quadrant maps' 8), j = 60 the zero vector; 61 <= j < 97 the first
    // 12 doubled on the far reference, j = 97 its zero vector.  A row scores the
    // 4 offsets dx0 .. dx0 + 3 with v_qsad_pk_u16_u8 on byte-aligned window
    // words and keeps those of the neighbourhood (the first 3; the zero vector:
    // the first) inside +-16: the candidate search's 290 points, same keys.
    if (cand) {
        const int j = tid >> 1, hf = tid & 1, far = j >= 61, u = far ? j - 61 : j;
        const bool zero = far ? u == 36 : u == 60;
        const int c = zero ? 0 : u / 3, sc = far ? 2 : 1;
        const int cq = c >= 12 && c < 20 ? c - 12 : 0;
        const int vx = c < 8 ? pr->cx[c < 8 ? c : 0] : c < 12 ? pr->kx[c < 12 ? c - 8 : 0] : pr->qx[cq >> 1][cq & 1];
        const int vy = c < 8 ? pr->cy[c < 8 ? c : 0] : c < 12 ? pr->ky[c < 12 ? c - 8 : 0] : pr->qy[cq >> 1][cq & 1];
        const int dy = zero ? 0 : sc * vy + u % 3 - 1, dx0 = zero ? 0 : sc * vx - 1, nv = zero ? 1 : 3;
        // in range: window columns WOFF + dx0 .. + 18 (>= 6, <= 58) and rows WOFF + dy .. + 15
        const bool live = j < 98 && !(far && nr < 2) && dy >= -16 && dy <= 16 && dx0 + nv - 1 >= -16 && dx0 <= 16;
        uint64_t acc[2] = {0, 0};
        if (live) {
            const int col = WOFF + dx0, sh = col & 3;
            const uint32_t *w = (const uint32_t *)(s.win[far] + (WOFF + dy + 8 * hf) * WIN) + (col >> 2);
#pragma unroll
            for (int y = 0; y < 8; ++y) {
                const uint32_t *r = w + y * (WIN / 4);
                const uint32_t r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4], r5 = r[5];
                const uint32_t s0 = __builtin_amdgcn_alignbyte(r1, r0, sh), s1 = __builtin_amdgcn_alignbyte(r2, r1, sh);
                const uint32_t s2 = __builtin_amdgcn_alignbyte(r3, r2, sh), s3 = __builtin_amdgcn_alignbyte(r4, r3, sh);
                const uint32_t s4 = __builtin_amdgcn_alignbyte(r5, r4, sh);
                const uint32_t *cr = s.cur32 + (8 * hf + y) * 4;
                acc[0] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)s1 << 32 | s0, cr[0], acc[0]);
                acc[0] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)s2 << 32 | s1, cr[1], acc[0]);
                acc[1] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)s3 << 32 | s2, cr[2], acc[1]);
                acc[1] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)s4 << 32 | s3, cr[3], acc[1]);
            }
        }
        // the partner's half: quadrants 0, 1 (top) | 2, 3 (bottom)
        const uint64_t o0 = __shfl_xor(acc[0], 1), o1 = __shfl_xor(acc[1], 1);
        const uint64_t qa[4] = {hf ? o0 : acc[0], hf ? o1 : acc[1], hf ? acc[0] : o0, hf ? acc[1] : o1};
        // this lane keeps offsets hf and hf + 2
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int o = hf + 2 * k, dx = dx0 + o;
            if (!live || o >= nv || dx < -16 || dx > 16) continue;
            const int pidx = (dy + 16) * 33 + dx + 16;
            const unsigned long long tie = ((unsigned long long)(abs(dx) + abs(dy)) << 16) | (unsigned)pidx;
            unsigned qs[4];
#pragma unroll
            for (int qd = 0; qd < 4; ++qd) qs[qd] = (unsigned)(qa[qd] >> (16 * o)) & 0xffffu;
            const unsigned long long k16 = ((unsigned long long)(qs[0] + qs[1] + qs[2] + qs[3]) << 32) | tie;
            if (!far) {
                best[0] = umin64(best[0], k16);
#pragma unroll
                for (int qd = 0; qd < 4; ++qd) best[2 + qd] = umin64(best[2 + qd], ((unsigned long long)qs[qd] << 32) | tie);
            } else {
                best[1] = umin64(best[1], k16);
            }
        }
    }
    // full search: an item is (reference, dy, 4 consecutive dx from an
    // aligned window word); v_qsad_pk_u16_u8 scores the 4 offsets of 4 source
    // bytes against 8 window bytes at once, 16-bit sums (a 16x16 SAD fits),
    // one accumulator per 8x8 quadrant.  Same keys, so the same minima.
    for (int it = tid; !cand && it < nr * 33 * 9; it += 256) {
        const int r = it / (33 * 9), rem = it % (33 * 9), dy = rem / 9 - 16, dx0 = 4 * (rem % 9) - 16;
        const uint32_t *wrow = (const uint32_t *)(s.win[r] + (WOFF + dy) * WIN) + ((WOFF + dx0) >> 2);
        uint64_t acc[4] = {0, 0, 0, 0};
#pragma unroll
        for (int y = 0; y < 16; ++y) {
            const uint32_t *w = wrow + y * (WIN / 4);
            const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4];
            const int q = (y >> 3) * 2;
            acc[q] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w1 << 32 | w0, s.cur32[y * 4 + 0], acc[q]);
            acc[q] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w2 << 32 | w1, s.cur32[y * 4 + 1], acc[q]);
            acc[q + 1] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w3 << 32 | w2, s.cur32[y * 4 + 2], acc[q + 1]);
            acc[q + 1] = __builtin_amdgcn_qsad_pk_u16_u8((uint64_t)w4 << 32 | w3, s.cur32[y * 4 + 3], acc[q + 1]);
        }
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const int dx = dx0 + o;
            if (dx > 16) continue;
            const int pidx = (dy + 16) * 33 + dx + 16;
            const unsigned long long tie = ((unsigned long long)(abs(dx) + abs(dy)) << 16) | (unsigned)pidx;
            unsigned qs[4];
#pragma unroll
            for (int qd = 0; qd < 4; ++qd) qs[qd] = (unsigned)(acc[qd] >> (16 * o)) & 0xffffu;
            const unsigned long long k16 = ((unsigned long long)(qs[0] + qs[1] + qs[2] + qs[3]) << 32) | tie;
            if (r == 0) {
                best[0] = umin64(best[0], k16);
#pragma unroll
                for (int qd = 0; qd < 4; ++qd) best[2 + qd] = umin64(best[2 + qd], ((unsigned long long)qs[qd] << 32) | tie);
            } else {
                best[1] = umin64(best[1], k16);
            }
        }
    }
    const int wv = tid >> 6, L = tid & 63;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        unsigned long long v = best[b];
        for (int o = 1; o < 64; o <<= 1) {
            unsigned long long u = __shfl_xor(v, o);
            v = umin64(v, u);
        }
        if (L == 0) s.key[b][wv] = v;
    }
    __syncthreads();
    int bmx[6], bmy[6], bbest[6];
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const unsigned long long k = umin64(umin64(s.key[b][0], s.key[b][1]), umin64(s.key[b][2], s.key[b][3]));
        const int pidx = (int)(k & 0xffff);
        bbest[b] = (int)(k >> 32);
        bmx[b] = (pidx % 33 - 16) * 4; bmy[b] = (pidx / 33 - 16) * 4;
    }
    refine_all(s, nr, bmx, bmy, bbest);
    if (tid == 0)
#pragma unroll
        for (int b = 0; b < 6; ++b) { s.res[b][0] = bmx[b]; s.res[b][1] = bmy[b]; s.res[b][2] = bbest[b]; }
    __syncthreads();
    if (tid == 0) {
        fdvg_metab &m = out[q.mb0 + i];
        for (int r = 0; r < 2; ++r) {
            const int b = (r == 1 && nr < 2) ? 0 : r;
            m.mv16[r][0] = (int16_t)s.res[b][0]; m.mv16[r][1] = (int16_t)s.res[b][1]; m.sad16[r] = s.res[b][2];
        }
        for (int qd = 0; qd < 4; ++qd) {
            m.mv8[qd][0] = (int16_t)s.res[2 + qd][0]; m.mv8[qd][1] = (int16_t)s.res[2 + qd][1];
            m.sad8[qd] = s.res[2 + qd][2];
        }
        m.flags = (pr && pr->pskip >= sthr) ? 1 : 0;
    }
}

// xgrid: the XCD-strip grid width (xcd_grid) of the largest sequence
extern "C" int fdvg_launch_metab(const fdvg_seq *seqs_d, int nseq, int xgrid, void *metab, const void *nn,
                                 float thr, float sthr, hipStream_t st) {
    hipLaunchKernelGGL(k_me_exhaustive, dim3(xgrid, nseq), dim3(256), 0, st, seqs_d, (fdvg_metab *)metab,
                       (const fdvg_nnout *)nn, thr, sthr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// fdvgpu_post.hip -- the block-parallel work after the MB wavefront:
//   * in-loop deblocking (fdv.h:3479-3636): all vertical 8x8-grid edges, then all
//     horizontal ones.  Edges within a pass touch disjoint samples (an edge writes
//     p0,q0 and reads p1,q1, and edges are 8 apart), so each pass is one launch
//     with a thread per edge segment;
//   * border replication of the new reference (plane_extend, fdv.h:1603);
//   * serialisation of the MB records into the 16 coded streams in raster order
//     (code_block4 / fdv_coeff8_encode / pframe_encode emission, fdv.h:5268-5277,
//     2197-2203, 6036-6160, 6180-6193): a count pass, a per-stream exclusive scan,
//     and a write pass.
#include <hip/hip_runtime.h>
#include "fdvgpu_common.h"

__device__ __constant__ uint8_t p_SCAN4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
__device__ __constant__ uint8_t p_SCAN8[64] = {
     0,  1,  8, 16,  9,  2,  3, 10, 17, 24, 32, 25, 18, 11,  4,  5,
    12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,  6,  7, 14, 21, 28,
    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
    58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// ------------------------------------------------------------------ deblock


__device__ __forceinline__ int clipb(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

__device__ __forceinline__ void edge(uint8_t *q, int st, int al, int be, int tc) {
    int p1 = q[-2 * st], p0 = q[-st], q0 = q[0], q1 = q[st];
    if (abs(p0 - q0) < al && abs(p1 - p0) < be && abs(q1 - q0) < be) {
        int d = ((q0 - p0) * 4 + (p1 - q1) + 4) >> 3;
        d = d < -tc ? -tc : (d > tc ? tc : d);
        q[-st] = (uint8_t)clipb(p0 + d);
        q[0] = (uint8_t)clipb(q0 - d);
    }
}

// pass 0: thread per (row, vertical edge); pass 1: thread per (horizontal edge, column).
extern "C" __global__ void __launch_bounds__(256)
k_deblock(const PlaneRef *__restrict__ planes, int pass) {
    const PlaneRef P = planes[blockIdx.y];
    const int al = 6 + P.qp, be = 2 + P.qp / 4, tc = 1 + P.qp / 10;
    const int ne_v = (P.w - 1) / 8;                 // vertical edges per row: x = 8, 16, .. < w
    const int ne_h = (P.h - 1) / 8;                 // horizontal edges
    const int n = pass == 0 ? P.h * (ne_v > 0 ? ne_v : 0) : ne_h * P.w;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (pass == 0) {
            const int y = i / ne_v, e = i % ne_v;
            edge(P.p + (ptrdiff_t)y * P.stride + 8 * (e + 1), 1, al, be, tc);
        } else {
            const int e = i / P.w, x = i % P.w;
            edge(P.p + (ptrdiff_t)(8 * (e + 1)) * P.stride + x, P.stride, al, be, tc);
        }
    }
}

// Replicate the outermost samples into the FDVG_BORDER-wide border.  Each border
// sample copies the interior sample at its clamped coordinate, so there is no
// ordering between rows and columns.
extern "C" __global__ void __launch_bounds__(256)
k_border(const PlaneRef *__restrict__ planes) {
    const PlaneRef P = planes[blockIdx.y];
    const int B = FDVG_BORDER, fw = P.w + 2 * B;
    const int ntb = B * fw;                 // one of the top/bottom bands
    const int n = 2 * ntb + P.h * 2 * B;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int x, y;
        if (i < ntb)            { y = -B + i / fw;   x = -B + i % fw; }
        else if (i < 2 * ntb)   { int j = i - ntb; y = P.h + j / fw; x = -B + j % fw; }
        else                    { int j = i - 2 * ntb; y = j / (2 * B); int k = j % (2 * B); x = k < B ? -B + k : P.w + (k - B); }
        const int sx = x < 0 ? 0 : (x >= P.w ? P.w - 1 : x);
        const int sy = y < 0 ? 0 : (y >= P.h ? P.h - 1 : y);
        P.p[(ptrdiff_t)y * P.stride + x] = P.p[(ptrdiff_t)sy * P.stride + sx];
    }
}

// Copy a padded plane's interior to a tight buffer (recon readback).
extern "C" __global__ void __launch_bounds__(256)
k_unpad(const PlaneRef *__restrict__ planes, uint8_t *const *__restrict__ dsts) {
    const PlaneRef P = planes[blockIdx.y];
    uint8_t *d = dsts[blockIdx.y];
    const int n = P.w * P.h;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        d[i] = P.p[(ptrdiff_t)(i / P.w) * P.stride + i % P.w];
}

// Tight I420 source -> padded planes with borders (a reference loaded from the host).
extern "C" __global__ void __launch_bounds__(256)
k_pad(const PlaneRef *__restrict__ planes, const uint8_t *const *__restrict__ srcs) {
    const PlaneRef P = planes[blockIdx.y];
    const uint8_t *s = srcs[blockIdx.y];
    const int B = FDVG_BORDER, fw = P.w + 2 * B, n = fw * (P.h + 2 * B);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int y = i / fw - B, x = i % fw - B;
        const int sx = x < 0 ? 0 : (x >= P.w ? P.w - 1 : x);
        const int sy = y < 0 ? 0 : (y >= P.h ? P.h - 1 : y);
        P.p[(ptrdiff_t)y * P.stride + x] = s[(size_t)sy * P.w + sx];
    }
}

// ------------------------------------------------------------------ symbols


__device__ __forceinline__ uint32_t zz(int v) { return ((uint32_t)v << 1) ^ (uint32_t)(v >> 31); }
__device__ __forceinline__ int leb_n(uint32_t u) {
    return u < (1u << 7) ? 1 : u < (1u << 14) ? 2 : u < (1u << 21) ? 3 : u < (1u << 28) ? 4 : 5;
}

// A byte sink that either counts (out == nullptr) or writes.
struct Sink {
    uint8_t *base[FDVG_NS];
    uint32_t n[FDVG_NS];
    uint32_t lim[FDVG_NS];         // writable bytes from base[k] (capacity guard)
    __device__ __forceinline__ void byte(int k, unsigned v) {
        if (base[k] && n[k] < lim[k]) base[k][n[k]] = (uint8_t)v;
        n[k]++;
    }
    __device__ __forceinline__ void leb(int k, uint32_t u) {
        while (u >= 0x80) { byte(k, 0x80 | (u & 0x7f)); u >>= 7; }
        byte(k, u);
    }
};

// Structure symbols and region / any-flags of one MB, in emission order
// (pframe_encode, fdv.h:6036-6160; code_residual flags 5385-5410; chroma_block
// any-flags 5599-5606).  Coefficient blocks are emitted by the lanes (below).
// The mode byte itself (to stream ctx, 0-3) is emitted by the caller, so that
// every stream index here is a constant and the sink stays in registers.
__device__ __forceinline__ void emit_structure(Sink &S, const fdvg_mbrec *R) {
    if (R->mode == 0) {
        if (R->ncand > 1) S.byte(SG_MERGE, R->merge_idx);
        return;
    }
    if (R->mode == 1) {
        S.byte(SG_REF, R->ref);
        S.leb(SG_MVX, zz(R->mvd[0][0]));
        S.leb(SG_MVY, zz(R->mvd[0][1]));
    } else if (R->mode == 2) {
        S.byte(SG_SUB, R->sub);
    } else {
        for (int q = 0; q < 4; ++q) { S.leb(SG_MVX, zz(R->mvd[q][0])); S.leb(SG_MVY, zz(R->mvd[q][1])); }
    }
    for (int r = 0; r < 4; ++r) S.byte(SG_FL, R->lflag[r]);
    if (R->ckind == 1 || R->ckind == 2)
        for (int p = 0; p < 2; ++p) S.byte(SG_FL, R->cflag[p]);
}

__device__ __forceinline__ int mb_ctx(const fdvg_mbrec *recs, const SymSeq &q, int mx, int my) {
    const int mb = q.mb0 + my * q.mw + mx;
    const int L = mx > 0 ? recs[mb - 1].mode : 0;
    const int A = my > 0 ? recs[mb - q.mw].mode : 0;
    return (L ? 2 : 0) + (A ? 1 : 0);
}

// The coefficient blocks of one MB as 30 unit slots, one per lane, numbered so
// that lane order is emission order within every stream:
//   0-15  luma 4x4 (region r = u/4, block u%4)          -> CN / CL   when lflag[r] == 1
//   16-19 luma 8x8 (region u-16)                         -> CN8 / CL8 when lflag[r] == 2
//   20-27 chroma 4x4 (plane (u-20)/4, block (u-20)%4)    -> CNC / CLC
//   28-29 chroma 8x8 (plane u-28, INTRA region flag 2)   -> CN8 / CL8
// Chroma 4x4 units are coded when the MB is INTER8 (no flag), when the INTER16
// any-flag is set, or when the INTRA region flag is 1 (fdv.h:5385-5410, 5599-5626).
__device__ __forceinline__ int unit_of(const fdvg_mbrec *R, int u, const int16_t **lev, int *n,
                                       int *cn, int *cl) {
    if (R->mode == 0 || u >= 30) return 0;
    if (u < 16) {
        if (R->lflag[u >> 2] != 1) return 0;
        *lev = R->lev + 64 * (u >> 2) + 16 * (u & 3); *n = 16; *cn = SG_CN; *cl = SG_CL; return 1;
    }
    if (u < 20) {
        if (R->lflag[u - 16] != 2) return 0;
        *lev = R->lev + 64 * (u - 16); *n = 64; *cn = SG_CN8; *cl = SG_CL8; return 1;
    }
    if (u < 28) {
        const int p = (u - 20) >> 2;
        const int on = R->ckind == 3 || (R->ckind == 1 && R->cflag[p]) || (R->ckind == 2 && R->cflag[p] == 1);
        if (!on) return 0;
        *lev = R->clev + 64 * p + 16 * ((u - 20) & 3); *n = 16; *cn = SG_CNC; *cl = SG_CLC; return 1;
    }
    if (R->ckind != 2 || R->cflag[u - 28] != 2) return 0;
    *lev = R->clev + 64 * (u - 28); *n = 64; *cn = SG_CN8; *cl = SG_CL8; return 1;
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
    const int L = threadIdx.x & 63;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(x, o, 64);
        if (L >= o) x += t;
    }
    return x - v;
}

// One wavefront per MB.  pass 0: per-MB byte counts into cnt[mb*16 + k];
// pass 1: write stream k of sequence s at compact + bases[s*16 + k] + cnt[mb*16 + k]
// (cnt holding the exclusive scan over the sequence's MBs by then).
extern "C" __global__ void __launch_bounds__(512)
k_sym(const SymSeq *__restrict__ seqs, const fdvg_mbrec *__restrict__ recs,
      uint32_t *__restrict__ cnt, int pass, uint8_t *__restrict__ compact,
      const uint32_t *__restrict__ bases, uint32_t cap) {
    const SymSeq q = seqs[blockIdx.y];
    const int nmb = q.mw * q.mh;
    const int L = threadIdx.x & 63;
    const int i = blockIdx.x * (int)(blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= nmb) return;
    const int mb = q.mb0 + i;
    // the MB's record into LDS first (16 bytes a lane): every later read of
    // levels and flags is then an LDS read, not a dependent global load
    __shared__ __align__(16) fdvg_mbrec rec_s[8];
    static_assert(sizeof(fdvg_mbrec) % 16 == 0 && sizeof(fdvg_mbrec) / 16 <= 64, "record staging");
    fdvg_mbrec *Rs = &rec_s[threadIdx.x >> 6];
    if (L < (int)(sizeof(fdvg_mbrec) / 16))
        ((uint4 *)Rs)[L] = ((const uint4 *)&recs[mb])[L];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const fdvg_mbrec *R = Rs;

    // this lane's coefficient unit: EOB count and LEB bytes of its levels
    const int16_t *lev = nullptr;
    int n = 0, cn = 0, cl = 0;
    const int on = unit_of(R, L, &lev, &n, &cn, &cl);
    int last = -1, lb = 0;
    if (on) {
        const uint8_t *scan = n == 16 ? p_SCAN4 : p_SCAN8;
        for (int k = 0; k < n; ++k) if (lev[scan[k]]) last = k;
        for (int k = 0; k <= last; ++k) lb += leb_n(zz(lev[scan[k]]));
    }
    // per coefficient-stream pair: lanes' byte counts and their in-MB offsets
    // (vector values: a lane indexes them with its own pair)
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    const int pr = !on ? -1 : cn == SG_CN ? 0 : cn == SG_CN8 ? 1 : 2;
    u4v offn = {0u, 0u, 0u, 0u}, offl = {0u, 0u, 0u, 0u}, totn = {0u, 0u, 0u, 0u}, totl = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t vn = pr == c ? 1u : 0u, vl = pr == c ? (uint32_t)lb : 0u;
        offn[c] = wave_excl_scan(vn);
        offl[c] = wave_excl_scan(vl);
        totn[c] = __shfl(offn[c] + vn, 63, 64);
        totl[c] = __shfl(offl[c] + vl, 63, 64);
    }

    if (pass == 1 && on) {
        // lanes write their blocks at their offsets inside the MB's run of the
        // two streams they feed (capacity-guarded like Sink::byte)
        const uint32_t bn = bases[blockIdx.y * FDVG_NS + cn], bl = bases[blockIdx.y * FDVG_NS + cl];
        const uint32_t limn = bn < cap ? cap - bn : 0, liml = bl < cap ? cap - bl : 0;
        uint32_t nn = cnt[(size_t)mb * FDVG_NS + cn] + offn[pr], nl = cnt[(size_t)mb * FDVG_NS + cl] + offl[pr];
        if (nn < limn) compact[bn + nn] = (uint8_t)(last + 1);
        const uint8_t *scan = n == 16 ? p_SCAN4 : p_SCAN8;
        for (int k = 0; k <= last; ++k) {
            uint32_t u = zz(lev[scan[k]]);
            while (u >= 0x80) { if (nl < liml) compact[bl + nl] = (uint8_t)(0x80 | (u & 0x7f)); ++nl; u >>= 7; }
            if (nl < liml) compact[bl + nl] = (uint8_t)u;
            ++nl;
        }
    }
    if (L == 0) {
        // the structure symbols, by one lane
        Sink S;
#pragma unroll
        for (int k = 0; k < FDVG_NS; ++k) {
            if (pass == 0) { S.base[k] = nullptr; S.n[k] = 0; S.lim[k] = 0; }
            else {
                const uint32_t b0 = bases[blockIdx.y * FDVG_NS + k];
                S.base[k] = compact + b0;
                S.n[k] = cnt[(size_t)mb * FDVG_NS + k];
                S.lim[k] = b0 < cap ? cap - b0 : 0;
            }
        }
        const int ctx = mb_ctx(recs, q, i % q.mw, i / q.mw);
        emit_structure(S, R);
        if (pass == 1) {
            // the mode byte, first in its MB's run of stream ctx (0-3)
            const uint32_t b0 = bases[blockIdx.y * FDVG_NS + ctx], n0 = cnt[(size_t)mb * FDVG_NS + ctx];
            if (b0 < cap && n0 < cap - b0) compact[b0 + n0] = R->mode;
        }
        if (pass == 0) {
            S.n[SG_CN] += totn[0]; S.n[SG_CL] += totl[0];
            S.n[SG_CN8] += totn[1]; S.n[SG_CL8] += totl[1];
            S.n[SG_CNC] += totn[2]; S.n[SG_CLC] += totl[2];
#pragma unroll
            for (int k = 0; k < FDVG_NS; ++k) cnt[(size_t)mb * FDVG_NS + k] = S.n[k] + (k == ctx ? 1u : 0u);
        }
    }
}

// Exclusive scan of one (sequence, stream) column of cnt[], in place; the total
// goes to totals[seq*16 + stream].  One block per column.
extern "C" __global__ void __launch_bounds__(256)
k_scan(const SymSeq *__restrict__ seqs, uint32_t *__restrict__ cnt, uint32_t *__restrict__ totals) {
    const SymSeq q = seqs[blockIdx.y];
    const int k = blockIdx.x;
    const int nmb = q.mw * q.mh;
    __shared__ uint32_t part[256];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nmb; base += 256) {
        const int i = base + threadIdx.x;
        const uint32_t v = i < nmb ? cnt[(size_t)(q.mb0 + i) * FDVG_NS + k] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            uint32_t t = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += t;
            __syncthreads();
        }
        const uint32_t incl = part[threadIdx.x];
        if (i < nmb) cnt[(size_t)(q.mb0 + i) * FDVG_NS + k] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 255) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.y * FDVG_NS + k] = carry;
}

extern "C" __global__ void k_zero_u32(uint32_t *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Start of every (sequence, stream) run in the compact output, and the total.
extern "C" __global__ void k_bases(const uint32_t *__restrict__ totals, uint32_t *__restrict__ bases, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < n; ++i) { bases[i] = run; run += totals[i]; }
        bases[n] = run;
    }
}

// ------------------------------------------------------------ host launchers

extern "C" int fdvg_launch_deblock(const void *planes, int nplanes, int maxpx, int bs, hipStream_t st) {
    const PlaneRef *P = (const PlaneRef *)planes;
    bs = bs == 64 || bs == 128 ? bs : 256;
    int gx = (maxpx / 8 + bs - 1) / bs;
    gx = gx < 1 ? 1 : gx > 4096 ? 4096 : gx;
    hipLaunchKernelGGL(k_deblock, dim3(gx, nplanes), dim3(bs), 0, st, P, 0);
    hipLaunchKernelGGL(k_deblock, dim3(gx, nplanes), dim3(bs), 0, st, P, 1);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvg_launch_border(const void *planes, int nplanes, int maxbpx, hipStream_t st) {
    int gx = (maxbpx + 255) / 256;
    gx = gx < 1 ? 1 : gx > 1024 ? 1024 : gx;
    hipLaunchKernelGGL(k_border, dim3(gx, nplanes), dim3(256), 0, st, (const PlaneRef *)planes);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvg_launch_pad(const void *planes, const uint8_t *const *srcs, int nplanes, int maxpx,
                               hipStream_t st) {
    int gx = (maxpx + 255) / 256;
    gx = gx < 1 ? 1 : gx > 2048 ? 2048 : gx;
    hipLaunchKernelGGL(k_pad, dim3(gx, nplanes), dim3(256), 0, st, (const PlaneRef *)planes, srcs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvg_launch_unpad(const void *planes, uint8_t *const *dsts, int nplanes, int maxpx,
                                 hipStream_t st) {
    int gx = (maxpx + 255) / 256;
    gx = gx < 1 ? 1 : gx > 2048 ? 2048 : gx;
    hipLaunchKernelGGL(k_unpad, dim3(gx, nplanes), dim3(256), 0, st, (const PlaneRef *)planes, dsts);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvg_launch_zero(uint32_t *p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvg_launch_symbols(const void *symseqs, int nseq, int maxmb, const fdvg_mbrec *recs,
                                   uint32_t *cnt, uint32_t *totals, uint32_t *bases, uint8_t *compact,
                                   size_t cap, int waves, hipStream_t st) {
    const uint32_t cap32 = cap > 0xffffffffu ? 0xffffffffu : (uint32_t)cap;
    const SymSeq *S = (const SymSeq *)symseqs;
    waves = waves == 1 || waves == 2 || waves == 8 ? waves : 4;    // MBs (wavefronts) per workgroup
    int gx = (maxmb + waves - 1) / waves;
    hipLaunchKernelGGL(k_sym, dim3(gx, nseq), dim3(64 * waves), 0, st, S, recs, cnt, 0, compact, (const uint32_t *)bases, cap32);
    hipLaunchKernelGGL(k_scan, dim3(FDVG_NS, nseq), dim3(256), 0, st, S, cnt, totals);
    hipLaunchKernelGGL(k_bases, dim3(1), dim3(64), 0, st, (const uint32_t *)totals, bases, nseq * FDVG_NS);
    hipLaunchKernelGGL(k_sym, dim3(gx, nseq), dim3(64 * waves), 0, st, S, recs, cnt, 1, compact, (const uint32_t *)bases, cap32);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Staging of a frame step's symbol streams for the pipelined interface: the
// stream bases (n + 1 words), the MB wavefront's timeout flag and an overflow
// flag into dbases, then exactly the bases[n] bytes of the compacted streams
// into dst (both in host memory mapped for the device), so that the host can
// read them once this kernel has completed while later frame steps already run
// on the stream.  Bytes beyond cap are not copied (the overflow flag says so).
extern "C" __global__ void __launch_bounds__(256)
k_stage_streams(const uint32_t *__restrict__ bases, const unsigned *__restrict__ ctl, int n,
                const uint8_t *__restrict__ src, size_t cap, uint32_t *__restrict__ dbases,
                uint8_t *__restrict__ dst) {
    const size_t total = bases[n];
    const size_t m = total < cap ? total : cap;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x, nt = (size_t)gridDim.x * blockDim.x;
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k <= n; k += blockDim.x) dbases[k] = bases[k];
        if (threadIdx.x == 0) { dbases[n + 1] = ctl[1]; dbases[n + 2] = total > cap ? 1u : 0u; }
    }
    // 16-byte chunks (both buffers are 16-byte aligned), then the tail
    const size_t nv = m / 16;
    for (size_t k = t0; k < nv; k += nt) ((uint4 *)dst)[k] = ((const uint4 *)src)[k];
    for (size_t k = nv * 16 + t0; k < m; k += nt) dst[k] = src[k];
}

extern "C" int fdvg_launch_stage(const uint32_t *bases, const unsigned *ctl, int n, const uint8_t *src, size_t cap,
                                 uint32_t *dbases, uint8_t *dst, hipStream_t st) {
    hipLaunchKernelGGL(k_stage_streams, dim3(64), dim3(256), 0, st, bases, ctl, n, src, cap, dbases, dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

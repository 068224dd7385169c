// fdvgpu_nn.hip -- the learned motion proposer: a small BF16 network on MFMA.
//
// Not part of the reference codec: it only chooses which whole-pel vectors the
// block-parallel search evaluates, and flags MBs it is confident are SKIP.  All
// codec arithmetic downstream stays integer and bit-exact.
//
// Input of one MB: its 16x16 luma block and the 64x64 co-located window of the
// nearest reference, both downsampled 2x (8x8 and 32x32, rounded 2x2 means).  A
// fixed correlation front end turns them into a 17x17 map over the half-res
// displacements -8..8 (+-16 whole pels): channel 0 = SAD relative to the
// quantiser step, SAD / (16 * 2^((qp - 4) / 6)) -- whether a residual survives
// quantisation is what SKIP turns on -- and channel 1 = (SAD - min SAD) / 256.  Then three 3x3 convolutions with 32 channels and ReLU
// (zero padding), and one linear layer applied at every position: output 0 is
// the position's vector logit, output 1 averaged over the positions (plus its
// bias) is the SKIP logit.
//
// Numerics: weights are rounded to bf16 for the forward pass; the inputs of the
// three convolutions (the features, then the two 32-channel activations) are
// bf16 -- the convolutions run on v_mfma_f32_32x32x16_bf16 with fp32
// accumulation -- and the linear layer runs on fp32 VALU with the bf16-rounded
// weights; biases are fp32.  The backward pass feeds MFMA with bf16 gradients
// and accumulates fp32.  Master weights and Adam state are fp32.
//
// Weight layout (floats): w1[32][18] (k = tap*2 + ci), b1[32], w2[32][288]
// (k = tap*32 + ci, tap = ky*3 + kx), b2[32], w3[32][288], b3[32], wf[2][32], bf[2].
#include <hip/hip_runtime.h>
#include "fdvgpu_common.h"
#include "fdvgpu_dev.h"

#define NN_G 17
#define NN_P 289
#define NN_PT 320                  // positions padded to ten 32-row tiles
#define NN_C 32
#define NN_K 288
#define NN_K1 18
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + 32 * 18)
#define OFF_W2 (OFF_B1 + 32)
#define OFF_B2 (OFF_W2 + 32 * 288)
#define OFF_W3 (OFF_B2 + 32)
#define OFF_B3 (OFF_W3 + 32 * 288)
#define OFF_WF (OFF_B3 + 32)
#define OFF_BF (OFF_WF + 64)
#define NN_NPARAM (OFF_BF + 2)

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

// One gfx950 32x32x16 bf16 MFMA.  Lane l holds row (A) / column (B) l % 32 and
// the eight k slots 8 * (l / 32) .. +7; the caller fills A and B with the same
// k in the same slot, so which slot holds which k is free.
__device__ __forceinline__ f16v mfma16(s8v a, s8v b, f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}
__device__ __forceinline__ s8v cat8(bf4 lo, bf4 hi) { return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7); }

__device__ __forceinline__ uint16_t f2bf(float f) {          // round to nearest even
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float rbf(float f) { return bf2f(f2bf(f)); }

// Position p = y*17 + x shifted by tap (ky-1, kx-1); -1 if outside the map.
__device__ __forceinline__ int nn_shift(int p, int tap) {
    if (p >= NN_P) return -1;
    const int y = p / NN_G + tap / 3 - 1, x = p % NN_G + tap % 3 - 1;
    return (y < 0 || y >= NN_G || x < 0 || x >= NN_G) ? -1 : y * NN_G + x;
}
// ... and by the opposite offset (the transposed convolution)
__device__ __forceinline__ int nn_unshift(int p, int tap) {
    if (p >= NN_P) return -1;
    const int y = p / NN_G - (tap / 3 - 1), x = p % NN_G - (tap % 3 - 1);
    return (y < 0 || y >= NN_G || x < 0 || x >= NN_G) ? -1 : y * NN_G + x;
}

struct __align__(16) NnLds {
    uint16_t w2[NN_C * NN_K], w3[NN_C * NN_K];   // bf16, [co][k]
    uint16_t w1[32 * 24];                        // bf16, [co][k], k = tap*2 + ci padded to 24
    float b1[32], b2[32], b3[32], wf[64], bf[2];
    float feat[NN_P * 2];                        // fp32 features, then ...
    uint16_t fb[NN_P * 2];                       // ... their bf16 copy, [pos][ci]
    uint16_t a1[NN_PT * NN_C], a2[NN_PT * NN_C]; // bf16 [pos][c], post-ReLU
    float a3[NN_PT * NN_C];                      // fp32 [pos][c], post-ReLU (reused by the backward pass)
    uint16_t dz[NN_PT * NN_C];                   // bf16 gradient buffer
    float lg[NN_PT], sk[NN_PT];
    float red[8][8];
    int ired[8][8];
};

// ---- inputs: the correlation map of one MB (luma of cur at (bx, by), ref plane)
__device__ void nn_features(const uint8_t *cur, int cw, const uint8_t *ref, int rs, int bx, int by, int qp,
                            float *feat, float *redf, int *redi) {
    __shared__ uint8_t c8[64], r32[32 * 32];
    __shared__ int cost[NN_P];
    const int t = threadIdx.x;
    if (t < 64) {
        const int y = t >> 3, x = t & 7;
        const uint8_t *p = cur + (size_t)(by + 2 * y) * cw + bx + 2 * x;
        c8[t] = (uint8_t)((p[0] + p[1] + p[cw] + p[cw + 1] + 2) >> 2);
    }
    for (int k = t; k < 1024; k += blockDim.x) {
        const int y = k >> 5, x = k & 31;
        const uint8_t *p = ref + (ptrdiff_t)(by - 24 + 2 * y) * rs + bx - 24 + 2 * x;
        r32[k] = (uint8_t)((p[0] + p[1] + p[rs] + p[rs + 1] + 2) >> 2);
    }
    __syncthreads();
    int mn = 1 << 30;
    for (int p = t; p < NN_P; p += blockDim.x) {
        const int dy = p / NN_G - 8, dx = p % NN_G - 8;
        int s = 0;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) s += abs((int)c8[y * 8 + x] - (int)r32[(12 + dy + y) * 32 + 12 + dx + x]);
        cost[p] = s;
        mn = min(mn, s);
    }
    for (int o = 32; o > 0; o >>= 1) mn = min(mn, __shfl_xor(mn, o));
    if ((t & 63) == 0) redi[t >> 6] = mn;
    __syncthreads();
    mn = redi[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) mn = min(mn, redi[w]);
    const float s0 = 1.0f / (16.0f * exp2f((float)(qp - 4) * (1.0f / 6.0f)));
    for (int p = t; p < NN_P; p += blockDim.x) {
        feat[2 * p] = (float)cost[p] * s0;
        feat[2 * p + 1] = (float)(cost[p] - mn) * (1.0f / 256.0f);
    }
    (void)redf;
    __syncthreads();
}

// ---- parameters into LDS (bf16-rounded weights, fp32 biases)
__device__ void nn_load(NnLds &S, const float *W) {
    const int t = threadIdx.x;
    for (int k = t; k < NN_C * NN_K; k += blockDim.x) { S.w2[k] = f2bf(W[OFF_W2 + k]); S.w3[k] = f2bf(W[OFF_W3 + k]); }
    for (int k = t; k < 32 * 24; k += blockDim.x) S.w1[k] = (k % 24) < 18 ? f2bf(W[OFF_W1 + (k / 24) * 18 + k % 24]) : 0;
    if (t < 32) { S.b1[t] = W[OFF_B1 + t]; S.b2[t] = W[OFF_B2 + t]; S.b3[t] = W[OFF_B3 + t]; }
    if (t < 64) S.wf[t] = rbf(W[OFF_WF + t]);
    if (t < 2) S.bf[t] = W[OFF_BF + t];
    __syncthreads();
}

// One 32x32 output tile of a 32-channel 3x3 convolution: rows = positions
// 32*tile .. +31, columns = output channels; K = 288 (tap*32 + ci).
__device__ __forceinline__ f16v conv_tile(const uint16_t *in, const uint16_t *w, int tile) {
    const int l = threadIdx.x & 63, i = l & 31, g = l >> 5;
    const int p = tile * 32 + i;
    f16v c = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int q = nn_shift(p, tap);
#pragma unroll
        for (int c0 = 0; c0 < 32; c0 += 16) {
            const int ci = c0 + 8 * g;
            bf4 a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
            if (q >= 0) { a0 = *(const bf4 *)&in[q * NN_C + ci]; a1 = *(const bf4 *)&in[q * NN_C + ci + 4]; }
            const uint16_t *wb = &w[i * NN_K + tap * 32 + ci];           // B[k][co = i] = w[co][k]
            c = mfma16(cat8(a0, a1), cat8(*(const bf4 *)wb, *(const bf4 *)(wb + 4)), c);
        }
    }
    return c;
}

// conv1 tile (2 -> 32 channels): K = 18 (k = tap*2 + ci) padded to 32; the
// eight k of a lane are four taps x two channels.
__device__ __forceinline__ f16v conv1_tile(const uint16_t *fb, const uint16_t *w1, int tile) {
    const int l = threadIdx.x & 63, i = l & 31, g = l >> 5;
    const int p = tile * 32 + i;
    f16v c = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 16) {
        const int kb = k0 + 8 * g;
        s8v a = {0, 0, 0, 0, 0, 0, 0, 0}, b = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int ta = (kb >> 1) + u, q = ta < 9 ? nn_shift(p, ta) : -1;
            if (q >= 0) { a[2 * u] = (short)fb[2 * q]; a[2 * u + 1] = (short)fb[2 * q + 1]; }
        }
        if (kb < 24) b = cat8(*(const bf4 *)&w1[i * 24 + kb], *(const bf4 *)&w1[i * 24 + kb + 4]);
        c = mfma16(a, b, c);
    }
    return c;
}

// The forward pass of one MB from S.feat: S.a1, S.a2 (bf16), S.a3 (fp32),
// S.lg (position logits), S.sk (per-position SKIP terms); returns the SKIP logit
// (uniform).
__device__ float nn_forward(NnLds &S) {
    const int t = threadIdx.x, wv = t >> 6, l = t & 63, nw = blockDim.x >> 6;
    for (int k = t; k < NN_P * 2; k += blockDim.x) S.fb[k] = f2bf(S.feat[k]);
    __syncthreads();
    // conv1 on MFMA: 2 -> 32 channels
    for (int tile = wv; tile < NN_PT / 32; tile += nw) {
        const f16v c = conv1_tile(S.fb, S.w1, tile);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = (r >> 2) * 8 + (l >> 5) * 4 + (r & 3), co = l & 31, p = tile * 32 + i;
            float v = p < NN_P ? c[r] + S.b1[co] : 0.f;
            S.a1[p * NN_C + co] = f2bf(v > 0.f ? v : 0.f);
        }
    }
    __syncthreads();
    // conv2, conv3 on MFMA
    for (int layer = 0; layer < 2; ++layer) {
        const uint16_t *in = layer ? S.a2 : S.a1;
        const uint16_t *w = layer ? S.w3 : S.w2;
        const float *b = layer ? S.b3 : S.b2;
        for (int tile = wv; tile < NN_PT / 32; tile += nw) {
            const f16v c = conv_tile(in, w, tile);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = (r >> 2) * 8 + (l >> 5) * 4 + (r & 3), co = l & 31, p = tile * 32 + i;
                float v = p < NN_P ? c[r] + b[co] : 0.f;
                v = v > 0.f ? v : 0.f;
                if (layer) S.a3[p * NN_C + co] = v;
                else S.a2[p * NN_C + co] = f2bf(v);
            }
        }
        __syncthreads();
    }
    // the linear layer at every position
    for (int p = t; p < NN_PT; p += blockDim.x) {
        float o0 = 0.f, o1 = 0.f;
        if (p < NN_P) {
            o0 = S.bf[0];
            for (int c = 0; c < NN_C; ++c) { o0 += S.a3[p * NN_C + c] * S.wf[c]; o1 += S.a3[p * NN_C + c] * S.wf[32 + c]; }
        }
        S.lg[p] = o0; S.sk[p] = o1;
    }
    __syncthreads();
    float s = 0.f;
    for (int p = t; p < NN_P; p += blockDim.x) s += S.sk[p];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (l == 0) S.red[0][wv] = s;
    __syncthreads();
    float tot = 0.f;
    for (int w2 = 0; w2 < nw; ++w2) tot += S.red[0][w2];
    __syncthreads();
    return tot * (1.0f / NN_P) + S.bf[1];
}

// Softmax statistics of S.lg over the positions: max and sum of exp (uniform).
__device__ void nn_softmax(NnLds &S, float *mx_out, float *sum_out) {
    const int t = threadIdx.x, wv = t >> 6, l = t & 63, nw = blockDim.x >> 6;
    float m = -3.0e38f;
    for (int p = t; p < NN_P; p += blockDim.x) m = fmaxf(m, S.lg[p]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (l == 0) S.red[1][wv] = m;
    __syncthreads();
    m = S.red[1][0];
    for (int w = 1; w < nw; ++w) m = fmaxf(m, S.red[1][w]);
    float s = 0.f;
    for (int p = t; p < NN_P; p += blockDim.x) s += expf(S.lg[p] - m);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (l == 0) S.red[2][wv] = s;
    __syncthreads();
    s = 0.f;
    for (int w = 0; w < nw; ++w) s += S.red[2][w];
    __syncthreads();
    *mx_out = m; *sum_out = s;
}

// ---------------------------------------------------------------- training

// Features of a batch of tuples: tuple j is the MB (bx, by) of a frame whose
// luma and padded reference luma are given per tuple.
extern "C" __global__ void __launch_bounds__(256)
k_nn_feat(const uint8_t *const *__restrict__ curs, const int *__restrict__ cw, const uint8_t *const *__restrict__ refs,
          const int *__restrict__ rs, const int2 *__restrict__ pos, const int *__restrict__ qps, float *__restrict__ feats) {
    __shared__ float redf[8];
    __shared__ int redi[8];
    __shared__ float fe[NN_P * 2];
    const int j = blockIdx.x;
    nn_features(curs[j], cw[j], refs[j], rs[j], pos[j].x, pos[j].y, qps[j], fe, redf, redi);
    for (int k = threadIdx.x; k < NN_P * 2; k += blockDim.x) feats[(size_t)j * NN_P * 2 + k] = fe[k];
}

// One training step's forward + backward over a batch: tuple j has features
// feats[j], vector target tgt[j] (position index, or -1 for none) and SKIP label
// skip[j].  The SKIP cross-entropy weighs false positives by wfp:
// -y log p - wfp (1 - y) log(1 - p).  Gradients (sums over the batch)
// accumulate into grad[]; into loss[]: [0] vector cross-entropy, [1] SKIP
// cross-entropy, [2] top-1 hits, [3] top-8 hits, [4] tuples with p >= thr,
// [5] of which labelled SKIP, [6] tuples whose top-8 probability >= thr,
// [7] of which with the target in the top 8.
extern "C" __global__ void __launch_bounds__(256)
k_nn_train(const float *__restrict__ W, const float *__restrict__ feats, const int *__restrict__ tgt,
           const uint8_t *__restrict__ skip, int n, float wfp, float thr, float *__restrict__ grad,
           float *__restrict__ loss) {
    extern __shared__ uint8_t dyn_nn[];
    NnLds &S = *(NnLds *)dyn_nn;
    const int t = threadIdx.x, wv = t >> 6, l = t & 63, nw = blockDim.x >> 6;
    nn_load(S, W);
    for (int j = blockIdx.x; j < n; j += gridDim.x) {
        for (int k = t; k < NN_P * 2; k += blockDim.x) S.feat[k] = feats[(size_t)j * NN_P * 2 + k];
        __syncthreads();
        const float skl = nn_forward(S);
        float m, se;
        nn_softmax(S, &m, &se);
        const int tg = tgt[j];
        const float y = skip[j] ? 1.f : 0.f;
        const float ps = 1.f / (1.f + expf(-skl));
        // losses and the rank of the target
        if (t == 0) {
            float top[8];
            for (int k = 0; k < 8; ++k) top[k] = -3.0e38f;
            for (int p = 0; p < NN_P; ++p) {
                float v = S.lg[p];
                for (int k = 0; k < 8; ++k) if (v > top[k]) { const float u = top[k]; top[k] = v; v = u; }
            }
            float conf = 0.f;
            for (int k = 0; k < 8; ++k) conf += expf(top[k] - m) / se;
            if (tg >= 0) {
                atomicAdd(&loss[0], -(S.lg[tg] - m - logf(se)));
                int above = 0;
                for (int p = 0; p < NN_P; ++p) above += S.lg[p] > S.lg[tg];
                atomicAdd(&loss[2], above == 0 ? 1.f : 0.f);
                atomicAdd(&loss[3], above < 8 ? 1.f : 0.f);
                if (conf >= thr) { atomicAdd(&loss[6], 1.f); atomicAdd(&loss[7], above < 8 ? 1.f : 0.f); }
            }
            atomicAdd(&loss[1], -(y * logf(fmaxf(ps, 1e-12f)) + wfp * (1.f - y) * logf(fmaxf(1.f - ps, 1e-12f))));
            if (ps >= thr) { atomicAdd(&loss[4], 1.f); atomicAdd(&loss[5], y); }
        }
        __syncthreads();                         // the logits are overwritten below
        // d logits, d skip
        const float dsk = (y > 0.5f ? ps - 1.f : wfp * ps) * (1.0f / NN_P);
        for (int p = t; p < NN_PT; p += blockDim.x) {
            float d = 0.f;
            if (p < NN_P && tg >= 0) d = expf(S.lg[p] - m) / se - (p == tg ? 1.f : 0.f);
            S.lg[p] = d;
        }
        __syncthreads();
        // linear layer: dwf, dbf, and dZ3 = dA3 * [a3 > 0] into S.dz (bf16)
        if (t < 64) {
            const int o = t >> 5, c = t & 31;
            float g = 0.f;
            for (int p = 0; p < NN_P; ++p) g += (o ? dsk : S.lg[p]) * S.a3[p * NN_C + c];
            atomicAdd(&grad[OFF_WF + t], g);
        } else if (t == 64) {
            float g = 0.f;
            for (int p = 0; p < NN_P; ++p) g += S.lg[p];
            atomicAdd(&grad[OFF_BF], g);
            atomicAdd(&grad[OFF_BF + 1], dsk * NN_P);
        }
        for (int e = t; e < NN_PT * NN_C; e += blockDim.x) {
            const int p = e >> 5, c = e & 31;
            float d = 0.f;
            if (p < NN_P && S.a3[e] > 0.f) d = S.lg[p] * S.wf[c] + dsk * S.wf[32 + c];
            S.dz[e] = f2bf(d);
        }
        __syncthreads();
        // conv3 then conv2: dW, db, and the input gradient (masked by ReLU)
        for (int layer = 1; layer >= 0; --layer) {
            const uint16_t *in = layer ? S.a2 : S.a1;      // this layer's input
            const uint16_t *w = layer ? S.w3 : S.w2;
            const int ow = layer ? OFF_W3 : OFF_W2, ob = layer ? OFF_B3 : OFF_B2;
            // dW[co][tap*32 + ci] = sum_p dz[p][co] * in[shift(p, tap)][ci]: M = co, N = ci, K = p per tap
            for (int tap = wv; tap < 9; tap += nw) {
                f16v c = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                const int i = l & 31, g = l >> 5;
                for (int p0 = 0; p0 < NN_PT; p0 += 16) {
                    s8v a, b;
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int p = p0 + 8 * g + u;
                        a[u] = (short)S.dz[p * NN_C + i];                       // A[co = i][p]
                        const int q = nn_shift(p, tap);
                        b[u] = q >= 0 ? (short)in[q * NN_C + i] : (short)0;      // B[p][ci = i]
                    }
                    c = mfma16(a, b, c);
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int co = (r >> 2) * 8 + (l >> 5) * 4 + (r & 3), ci = l & 31;
                    atomicAdd(&grad[ow + co * NN_K + tap * 32 + ci], c[r]);
                }
            }
            if (t < 32) {
                float g = 0.f;
                for (int p = 0; p < NN_P; ++p) g += bf2f(S.dz[p * NN_C + t]);
                atomicAdd(&grad[ob + t], g);
            }
            // dIn[p][ci] = sum_tap,co dz[unshift(p, tap)][co] * w[co][tap*32 + ci]: M = p, N = ci, K = (tap, co)
            for (int tile = wv; tile < NN_PT / 32; tile += nw) {
                f16v c = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                const int i = l & 31, g = l >> 5, p = tile * 32 + i;
                for (int tap = 0; tap < 9; ++tap) {
                    const int q = nn_unshift(p, tap);
#pragma unroll
                    for (int c0 = 0; c0 < 32; c0 += 16) {
                        const int co = c0 + 8 * g;
                        bf4 a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
                        if (q >= 0) { a0 = *(const bf4 *)&S.dz[q * NN_C + co]; a1 = *(const bf4 *)&S.dz[q * NN_C + co + 4]; }
                        s8v b;
#pragma unroll
                        for (int u = 0; u < 8; ++u) b[u] = (short)w[(co + u) * NN_K + tap * 32 + i];
                        c = mfma16(cat8(a0, a1), b, c);
                    }
                }
                // masked by this layer's input activation; stage in a3 (fp32) until dz is free
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int pi = (r >> 2) * 8 + (l >> 5) * 4 + (r & 3), ci = l & 31, pp = tile * 32 + pi;
                    const float act = bf2f(in[pp * NN_C + ci]);
                    S.a3[pp * NN_C + ci] = (pp < NN_P && act > 0.f) ? c[r] : 0.f;
                }
            }
            __syncthreads();
            for (int e = t; e < NN_PT * NN_C; e += blockDim.x) S.dz[e] = f2bf(S.a3[e]);
            __syncthreads();
        }
        // conv1: dW1[co][tap*2 + ci] = sum_p dz[p][co] * feat[shift(p, tap)][ci], db1
        for (int e = t; e < 32 * 18 + 32; e += blockDim.x) {
            float g = 0.f;
            if (e < 32 * 18) {
                const int co = e / 18, k = e % 18, tap = k >> 1, ci = k & 1;
                for (int p = 0; p < NN_P; ++p) {
                    const int q = nn_shift(p, tap);
                    if (q >= 0) g += bf2f(S.dz[p * NN_C + co]) * bf2f(S.fb[2 * q + ci]);
                }
                atomicAdd(&grad[OFF_W1 + e], g);
            } else {
                const int co = e - 32 * 18;
                for (int p = 0; p < NN_P; ++p) g += bf2f(S.dz[p * NN_C + co]);
                atomicAdd(&grad[OFF_B1 + co], g);
            }
        }
        __syncthreads();
    }
}

// Adam on the fp32 master weights with the batch-mean gradient, then clear it.
extern "C" __global__ void k_nn_adam(float *__restrict__ W, float *__restrict__ grad, float *__restrict__ m1,
                                     float *__restrict__ m2, int n, float lr, float b1, float b2, float eps,
                                     float inv_batch, float c1, float c2) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const float g = grad[k] * inv_batch;
        const float a = b1 * m1[k] + (1.f - b1) * g, b = b2 * m2[k] + (1.f - b2) * g * g;
        m1[k] = a; m2[k] = b;
        W[k] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
        grad[k] = 0.f;
    }
}

// --------------------------------------------------------------- inference


// The inference path: one wavefront per MB, four per workgroup, no
// workgroup barriers after the weights are staged.
//  - The convolutions run on v_mfma_f32_32x32x16_bf16 in transposed form: a
//    tile is (output channels) x (32 positions), weights as the A operand and
//    activations as the B operand, so that a lane's accumulators hold 16
//    channels of one position (lane l: position 32 t + l % 32, channels
//    8 q + 4 (l / 32) + e for accumulator 4 q + e).  The epilogues then store
//    four channels (8 bytes) at a time and the linear layer is an in-lane sum
//    plus one exchange between the wave's halves.
//  - conv2 and conv3: K = 16 (half a tap's 32 input channels) per step, 18
//    steps.  All ten 32-position tiles accumulate at once (160 accumulator
//    registers), so every weight operand is read once per step and the wave's
//    single activation buffer is rewritten in place after the last step.
//  - Weight operands: both layers' weights in bf16, stored in LDS in lane
//    order (step, lane) -> 16 bytes, shared by the workgroup's four waves.
//  - Activation operands: rows of NI_RS bf16 (80 bytes: eight consecutive rows
//    cover the 32 banks once for 16-byte reads); the zero rows and column of
//    the padded grid are the 3x3 taps' padding.
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
#define NI_RS 40
#define NI_WAVES 4
// The map's positions sit in a zero-bordered grid of rows NI_PW wide (one zero
// column, shared by the right edge of a row and the left edge of the next), so
// that every 3x3 tap of a position is a constant offset from it and lands on a
// zero outside the map: position p at ni_pad(p); entries below NI_P0, at or
// above NI_P0 + 17 NI_PW, and in column 17 are zero.  NI_PR covers the last
// position's taps.
#define NI_PW 18
#define NI_P0 19
#define NI_PR (NI_P0 + 18 * NI_PW)        // 343: past the bottom zero row
__host__ __device__ constexpr int ni_pad(int p) { return (p / NN_G) * NI_PW + p % NN_G + NI_P0; }
__device__ __forceinline__ bool ni_zero(int e) { return e < NI_P0 || e >= NI_P0 + NN_G * NI_PW || (e - NI_P0) % NI_PW == NN_G; }
static_assert(ni_pad(NN_P - 1) + NI_PW + 1 < NI_PR && ni_pad(0) - NI_PW - 1 >= 0, "every tap inside the grid");
struct __align__(16) NiWave {
    uint16_t act[NI_PR * NI_RS];           // [padded pos][channel]; rows 0 .. NI_P0 - 1 also the features' scratch
    uint32_t fb[NI_PR];                    // the two bf16 features of each padded position
    int pad_[1];
};
struct __align__(16) NiShared {
    u4v wl[2][18][64];                     // conv2 / conv3 weight operands, bf16 x 8 per (step, lane)
    u4v w1[2][64];                         // conv1's, per K step and lane
    float b1[32], b2[32], b3[32], wf[64], bf[2];
    NiWave wave[NI_WAVES];
};

// Four channels' ReLU(acc + bias) as bf16 (round to nearest even), 8 bytes.
__device__ __forceinline__ uint2 ni_relu4(float a0, float a1, float a2, float a3, float4 b) {
    typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
    float v0 = a0 + b.x, v1 = a1 + b.y, v2 = a2 + b.z, v3 = a3 + b.w;
    v0 = v0 > 0.f ? v0 : 0.f; v1 = v1 > 0.f ? v1 : 0.f; v2 = v2 > 0.f ? v2 : 0.f; v3 = v3 > 0.f ? v3 : 0.f;
    const bf2 lo = {(__bf16)v0, (__bf16)v1}, hi = {(__bf16)v2, (__bf16)v3};
    return make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
}

// A transposed tile's epilogue: this lane's position (padded index pad, inside
// the map when live) gets its 16 channels' activations, four at a time.
__device__ __forceinline__ void ni_put(uint16_t *act, int pad, bool live, const f16v &c, const float *bias, int g) {
    if (!live) return;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        *(uint2 *)&act[pad * NI_RS + 8 * q + 4 * g] =
            ni_relu4(c[4 * q], c[4 * q + 1], c[4 * q + 2], c[4 * q + 3], *(const float4 *)&bias[8 * q + 4 * g]);
}

__device__ __forceinline__ u4v ni_pack8(const float *f) {
    u4v r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = (uint32_t)f2bf(f[2 * j]) | ((uint32_t)f2bf(f[2 * j + 1]) << 16);
    return r;
}

// The wavefront's smallest unsigned value: inside each 16-lane row by DPP (quad
// swaps, then rotations by 4 and 8), then the four row results by readlane; the
// result is uniform.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t k) {
    k = min(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0xB1, 0xf, 0xf, false));   // quad_perm [1,0,3,2]
    k = min(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x4E, 0xf, 0xf, false));   // quad_perm [2,3,0,1]
    k = min(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x124, 0xf, 0xf, false));  // row_ror:4
    k = min(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x128, 0xf, 0xf, false));  // row_ror:8
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)k, 0), b = (uint32_t)__builtin_amdgcn_readlane((int)k, 16);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)k, 32), d = (uint32_t)__builtin_amdgcn_readlane((int)k, 48);
    return min(min(a, b), min(c, d));
}

// The correlation map of one MB into w.fb (bf16 pairs), by one wavefront.
// Integer SADs and the fp32 feature formulas are those of nn_features.  Also
// the map's 4 lowest-cost positions into o->kx, o->ky (whole pels; ties to the
// smaller |dx| + |dy|, then the lower position): classical coarse search
// candidates beside the network's; and each 8x8 quadrant's 2 lowest-cost
// positions of its own map into o->qx, o->qy (same ties).
__device__ __forceinline__ void ni_features(NiWave &w, const uint8_t *cur, int cw, const uint8_t *ref, int rs,
                                            int bx, int by, int qp, fdvg_nnout *o) {
    const int l = threadIdx.x & 63;
    uint8_t *c8 = (uint8_t *)w.act;
    uint8_t *r32 = c8 + 64;
    int *cost = (int *)(c8 + 1088);
    {
        const int y = l >> 3, x = l & 7;
        const gmem_t<uint8_t> *p = gmem(cur) + (size_t)(by + 2 * y) * cw + bx + 2 * x;
        c8[l] = (uint8_t)((p[0] + p[1] + p[cw] + p[cw + 1] + 2) >> 2);
    }
    {
        // lane: r32 row l/2, columns 16 (l&1) .. +15, from two reference rows of 32 bytes (8-byte aligned)
        const int y = l >> 1, h = l & 1;
        const gmem_t<uint8_t> *p0 = gmem(ref) + (ptrdiff_t)(by - 24 + 2 * y) * rs + bx - 24 + 32 * h;
        const gmem_t<uint8_t> *p1 = p0 + rs;
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t a64 = *(const gmem_t<uint64_t> *)(p0 + 8 * k), b64 = *(const gmem_t<uint64_t> *)(p1 + 8 * k);
            const uint2 a = make_uint2((uint32_t)a64, (uint32_t)(a64 >> 32)), b = make_uint2((uint32_t)b64, (uint32_t)(b64 >> 32));
            uint32_t v = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t wa = e < 2 ? a.x : a.y, wb = e < 2 ? b.x : b.y;
                const int sh = (e & 1) * 16;
                const uint32_t s4 = ((wa >> sh) & 255) + ((wa >> (sh + 8)) & 255) + ((wb >> sh) & 255) +
                                    ((wb >> (sh + 8)) & 255) + 2;
                v |= ((s4 >> 2) & 255) << (8 * e);
            }
            o[k] = v;
        }
        *(uint4 *)(r32 + y * 32 + 16 * h) = make_uint4(o[0], o[1], o[2], o[3]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t cr[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) cr[k] = *(const uint32_t *)(c8 + 4 * k);
    int mn = 1 << 30;
    int cv[5];
    uint32_t qkey[4][5];                 // per quadrant (raster) and position: (cost, |dx| + |dy|, position)
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const int p = l + 64 * m;
        int sad = 1 << 30;
        if (p < NN_P) {
            const int dy = p / NN_G - 8, dx = p % NN_G - 8, c0 = 12 + dx, sh = c0 & 3;
            const uint8_t *row = r32 + (12 + dy) * 32 + (c0 & ~3);
            uint32_t aq[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int y = 0; y < 8; ++y) {
                const uint32_t d0 = *(const uint32_t *)(row + 32 * y), d1 = *(const uint32_t *)(row + 32 * y + 4),
                               d2 = *(const uint32_t *)(row + 32 * y + 8);
                const int qr = (y >> 2) * 2;
                aq[qr] = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d1, d0, sh), cr[2 * y], aq[qr]);
                aq[qr + 1] = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d2, d1, sh), cr[2 * y + 1], aq[qr + 1]);
            }
            sad = (int)(aq[0] + aq[1] + aq[2] + aq[3]);
            mn = min(mn, sad);
            const uint32_t tie = ((uint32_t)(abs(dx) + abs(dy)) << 9) | (uint32_t)p;
#pragma unroll
            for (int q = 0; q < 4; ++q) qkey[q][m] = (aq[q] << 14) | tie;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) qkey[q][m] = 0xFFFFFFFFu;
        }
        cv[m] = sad;
    }
    // each quadrant's 2 lowest-cost positions (a quadrant's cost < 2^12)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            uint32_t b = qkey[q][0];
#pragma unroll
            for (int m = 1; m < 5; ++m) b = min(b, qkey[q][m]);
            b = wave_min_u32(b);
#pragma unroll
            for (int m = 0; m < 5; ++m) if (qkey[q][m] == b) qkey[q][m] = 0xFFFFFFFFu;
            const int p = (int)(b & 511u);
            if (l == 0) { o->qx[q][k] = (int8_t)(2 * (p % NN_G - 8)); o->qy[q][k] = (int8_t)(2 * (p / NN_G - 8)); }
        }
    }
    {
        // keys (cost, |dx| + |dy|, position): cost < 2^14, distance < 2^5, position < 2^9
        uint32_t key[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const int p = l + 64 * m, dy = p / NN_G - 8, dx = p % NN_G - 8;
            key[m] = p < NN_P ? ((uint32_t)cv[m] << 14) | ((uint32_t)(abs(dx) + abs(dy)) << 9) | (uint32_t)p
                              : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t b = key[0];
#pragma unroll
            for (int m = 1; m < 5; ++m) b = min(b, key[m]);
            b = wave_min_u32(b);
#pragma unroll
            for (int m = 0; m < 5; ++m) if (key[m] == b) key[m] = 0xFFFFFFFFu;
            const int p = (int)(b & 511u);
            if (l == 0) { o->kx[k] = (int8_t)(2 * (p % NN_G - 8)); o->ky[k] = (int8_t)(2 * (p / NN_G - 8)); }
        }
    }
    mn = (int)wave_min_u32((uint32_t)mn);
    (void)cost;
    const float s0 = 1.0f / (16.0f * exp2f((float)(qp - 4) * (1.0f / 6.0f)));
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const int p = l + 64 * m;
        if (p < NN_P)
            w.fb[ni_pad(p)] = (uint32_t)f2bf((float)cv[m] * s0) | ((uint32_t)f2bf((float)(cv[m] - mn) * (1.0f / 256.0f)) << 16);
    }
}

// The wavefront's largest 64-bit key: inside each 16-lane row by DPP (quad
// swaps, then rotations by 4 and 8), then the four row results by readlane;
// the result is uniform.
#define NN_DPP_MAX(ctrl)                                                                                \
    do {                                                                                                \
        const uint32_t lo_ = __builtin_amdgcn_update_dpp(0u, (uint32_t)k, ctrl, 0xf, 0xf, false);        \
        const uint32_t hi_ = __builtin_amdgcn_update_dpp(0u, (uint32_t)(k >> 32), ctrl, 0xf, 0xf, false); \
        const uint64_t o_ = ((uint64_t)hi_ << 32) | lo_;                                               \
        k = o_ > k ? o_ : k;                                                                            \
    } while (0)
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t k) {
    NN_DPP_MAX(0xB1);              // quad_perm [1,0,3,2]
    NN_DPP_MAX(0x4E);              // quad_perm [2,3,0,1]
    NN_DPP_MAX(0x124);             // row_ror:4
    NN_DPP_MAX(0x128);             // row_ror:8
    uint64_t b = 0;
#pragma unroll
    for (int r = 0; r < 64; r += 16) {
        const uint64_t o = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(k >> 32), r) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, r);
        b = o > b ? o : b;
    }
    return b;
}
#undef NN_DPP_MAX

// Per MB of every sequence: the 8 most likely whole-pel vectors (nearest
// reference), their total probability and the SKIP probability.
extern "C" __global__ void __launch_bounds__(256, 1)
k_nn_infer(const fdvg_seq *__restrict__ seqs, int nseq, int total, const float *__restrict__ W,
           fdvg_nnout *__restrict__ out) {
    extern __shared__ __align__(16) uint8_t dyn_nn[];
    NiShared &S = *(NiShared *)dyn_nn;
    const int t = threadIdx.x, wv = t >> 6, l = t & 63, i32 = l & 31, g = l >> 5;
    NiWave &Wv = S.wave[wv];
    // stage the weights: B[k][co] for k = 16 s + 8 (lane / 32) + j, co = lane % 32
    for (int e = t; e < 2 * 18 * 64; e += blockDim.x) {
        const int layer = e / (18 * 64), st = (e / 64) % 18, ln = e % 64;
        const float *src = W + (layer ? OFF_W3 : OFF_W2) + (ln & 31) * NN_K + 16 * st + 8 * (ln >> 5);
        float f[8];
        const float4 x0 = *(const float4 *)src, x1 = *(const float4 *)(src + 4);
        f[0] = x0.x; f[1] = x0.y; f[2] = x0.z; f[3] = x0.w; f[4] = x1.x; f[5] = x1.y; f[6] = x1.z; f[7] = x1.w;
        S.wl[layer][st][ln] = ni_pack8(f);
    }
    if (t < 32) { S.b1[t] = W[OFF_B1 + t]; S.b2[t] = W[OFF_B2 + t]; S.b3[t] = W[OFF_B3 + t]; }
    if (t < 64) S.wf[t] = rbf(W[OFF_WF + t]);
    if (t < 2) S.bf[t] = W[OFF_BF + t];
    // this wave's zero rows
    for (int e = l; e < NI_PR; e += 64)
        if (ni_zero(e)) {
            Wv.fb[e] = 0;
#pragma unroll
            for (int c = 0; c < NI_RS / 2; ++c) ((uint32_t *)&Wv.act[e * NI_RS])[c] = 0;
        }
    // conv1's weight operands: k = 16 s + 8 g + j (k = tap*2 + ci, zero beyond 18)
    if (wv < 2) {
        float f[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 16 * wv + 8 * g + j;
            f[j] = k < NN_K1 ? W[OFF_W1 + i32 * NN_K1 + k] : 0.f;
        }
        S.w1[wv][l] = ni_pack8(f);
    }
    // per tile, this lane's padded position (recomputed where used rather than
    // held across the MB loop): the last tile's lanes past the map take
    // position 0, their outputs are discarded
    auto pb = [&](int tl) {
        const int p = tl * 32 + i32;
        return p < NN_P ? ni_pad(p) : ni_pad(0);
    };
    __syncthreads();
    const uint8_t *actb = (const uint8_t *)Wv.act;
    for (int gi = blockIdx.x * NI_WAVES + wv; gi < total; gi += gridDim.x * NI_WAVES) {
        int k = 0;
        while (k + 1 < nseq && seqs[k + 1].mb0 <= gi) ++k;
        const fdvg_seq &q = seqs[k];
        const int i = gi - q.mb0, mw = q.w >> 4;
        ni_features(Wv, q.cy, q.w, q.ry[0], q.ys, (i % mw) * 16, (i / mw) * 16, q.qp, &out[gi]);
        // the features used the first rows as scratch: the zero rows below NI_P0
        // are zeroed again; the rest it reaches are positions 0 .. 16 of the
        // first map row, which conv1 rewrites (the next zero entry is NI_P0 + 17)
        static_assert(64 + 1024 + 4 * NN_P <= (NI_P0 + NN_G) * NI_RS * 2, "features scratch short of the next zero entry");
        for (int e = l; e < NI_P0 * NI_RS / 2; e += 64) ((uint32_t *)Wv.act)[e] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // conv1: two K steps (taps 0-3 | 4-7 by lane half, then tap 8)
        {
#pragma unroll
            for (int tl = 0; tl < 10; ++tl) {
                f16v c = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                u4v a0, a1 = {0u, 0u, 0u, 0u};
                const int pt = pb(tl);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    // the lane half's four conv1 taps 4 g + j as padded offsets
                    const int tap = 4 * g + j;
                    a0[j] = Wv.fb[pt + (tap / 3 - 1) * NI_PW + tap % 3 - 1];
                }
                if (g == 0) a1[0] = Wv.fb[pt + NI_PW + 1];
                c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, S.w1[0][l]), __builtin_bit_cast(bf8, a0), c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, S.w1[1][l]), __builtin_bit_cast(bf8, a1), c, 0, 0, 0);
                ni_put(Wv.act, pt, tl * 32 + i32 < NN_P, c, S.b1, g);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        f16v acc[10];
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
#pragma unroll
            for (int tl = 0; tl < 10; ++tl)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[tl][r] = 0.f;
#pragma unroll
            for (int st = 0; st < 18; ++st) {
                const int tap = st >> 1;
                const int sh = (tap / 3 - 1) * NI_PW + tap % 3 - 1;
                const bf8 b = __builtin_bit_cast(bf8, S.wl[layer][st][l]);
                // the step's ten A operands are read before its MFMAs, so that
                // the reads are in flight together (a per-tile base one row and
                // one position back keeps every tap's offset a non-negative
                // immediate)
                u4v av[10];
#pragma unroll
                for (int tl = 0; tl < 10; ++tl) {
                    const int off = (pb(tl) - NI_PW - 1) * (NI_RS * 2) + 16 * g + (sh + NI_PW + 1) * (NI_RS * 2) + 32 * (st & 1);
                    av[tl] = *(const u4v *)(actb + off);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int tl = 0; tl < 10; ++tl)
                    acc[tl] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, __builtin_bit_cast(bf8, av[tl]), acc[tl], 0, 0, 0);
            }
            if (layer == 0) {
                // every read of the buffer has completed (its values fed the MFMAs): rewrite it with conv2
#pragma unroll
                for (int tl = 0; tl < 10; ++tl) ni_put(Wv.act, pb(tl), tl * 32 + i32 < NN_P, acc[tl], S.b2, g);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        // conv3 epilogue + linear layer: a lane sums its 16 channels of its
        // position, the two halves of the wave (the other 16 channels) are added
        // by one exchange; every position's logit is then on lanes i and i + 32.
        float skacc = 0.f;
        float lg[10];
#pragma unroll
        for (int tl = 0; tl < 10; ++tl) {
            const bool live = tl * 32 + i32 < NN_P;
            float lp = 0.f, sp = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // this lane's channels 8 q + 4 g .. +3 (16-byte LDS reads, the same for a half-wave)
                const float4 b = *(const float4 *)&S.b3[8 * q + 4 * g], w0 = *(const float4 *)&S.wf[8 * q + 4 * g],
                             w1 = *(const float4 *)&S.wf[32 + 8 * q + 4 * g];
                const float bb[4] = {b.x, b.y, b.z, b.w}, ww0[4] = {w0.x, w0.y, w0.z, w0.w},
                            ww1[4] = {w1.x, w1.y, w1.z, w1.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float a = acc[tl][4 * q + e] + bb[e];
                    a = (live && a > 0.f) ? a : 0.f;
                    lp += a * ww0[e];
                    sp += a * ww1[e];
                }
            }
            skacc += sp;
            lp += __shfl_xor(lp, 32);
            lg[tl] = live ? lp + S.bf[0] : -3.0e38f;
        }
        for (int o = 32; o > 0; o >>= 1) skacc += __shfl_xor(skacc, o);
        const float skl = skacc * (1.0f / NN_P) + S.bf[1];
        // softmax statistics (each position counted once: the first half-wave)
        float m = -3.0e38f;
#pragma unroll
        for (int tl = 0; tl < 10; ++tl) m = fmaxf(m, lg[tl]);
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        float se = 0.f;
        if (g == 0) {
#pragma unroll
            for (int tl = 0; tl < 10; ++tl) se += lg[tl] > -1.0e38f ? expf(lg[tl] - m) : 0.f;
        }
        for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o);
        // top 8, ties to the lower position.  Each logit becomes a key: its bits
        // mapped to an unsigned order (-0 taken as +0, which compares equal to
        // it) above the position's complement, so the largest key is the largest
        // logit, the lowest position among equal ones; a taken or absent
        // position's key is 0, below every logit's.
        fdvg_nnout &o = out[gi];
        float conf = 0.f;
        uint64_t key[10];
#pragma unroll
        for (int tl = 0; tl < 10; ++tl) {
            const int pp = tl * 32 + i32;
            uint32_t u = __float_as_uint(lg[tl] + 0.0f);
            u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
            key[tl] = g == 0 && pp < NN_P ? ((uint64_t)u << 32) | (uint32_t)(0xffff - pp) : 0ull;
        }
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            uint64_t b = key[0];
#pragma unroll
            for (int tl = 1; tl < 10; ++tl) b = key[tl] > b ? key[tl] : b;
            b = wave_max_u64(b);
#pragma unroll
            for (int tl = 0; tl < 10; ++tl) key[tl] = key[tl] == b ? 0ull : key[tl];
            const uint32_t u = (uint32_t)(b >> 32);
            const float bv = __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
            const int bp = 0xffff - (int)(b & 0xffffu);
            conf += expf(bv - m) / se;
            if (l == 0) { o.cx[kk] = (int8_t)(2 * (bp % NN_G - 8)); o.cy[kk] = (int8_t)(2 * (bp / NN_G - 8)); }
        }
        if (l == 0) { o.conf = conf; o.pskip = 1.f / (1.f + expf(-skl)); }
    }
}

extern "C" size_t fdvg_nn_lds(void) { return sizeof(NnLds); }
extern "C" int fdvg_nn_nparam(void) { return NN_NPARAM; }
extern "C" int fdvgpu_nn_nparam(void) { return NN_NPARAM; }

extern "C" int fdvg_nn_setup(void) {
    return hipFuncSetAttribute((const void *)k_nn_infer, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sizeof(NiShared)) == hipSuccess &&
           hipFuncSetAttribute((const void *)k_nn_train, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)sizeof(NnLds)) == hipSuccess ? 0 : -1;
}

// wpc: proposer workgroups per CU (1 or 2; one fits a CU at a time, so 2 only
// shortens each workgroup's share).  Waves take the MBs of all sequences in
// turn (sequence k owns MBs mb0 .. mb0 + its MB count - 1, in order).
extern "C" int fdvg_launch_nn_infer(const fdvg_seq *seqs, int nseq, int total, const float *W, fdvg_nnout *out,
                                    int wpc, hipStream_t st) {
    static bool attr = false;
    if (!attr) {
        if (hipFuncSetAttribute((const void *)k_nn_infer, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sizeof(NiShared)) != hipSuccess) return -1;
        attr = true;
    }
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    }
    if (total <= 0) return 0;
    int gx = ncu * (wpc == 2 ? 2 : 1);
    const int need = (total + NI_WAVES - 1) / NI_WAVES;
    gx = gx > need ? need : gx;
    hipLaunchKernelGGL(k_nn_infer, dim3(gx), dim3(64 * NI_WAVES), sizeof(NiShared), st, seqs, nseq, total, W, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------- C ABI (training)

extern "C" int fdvgpu_nn_features(const uint8_t *const *curs, const int *cw, const uint8_t *const *refs,
                                  const int *rs, const int *pos, const int *qps, int n, float *feats, void *stream) {
    hipLaunchKernelGGL(k_nn_feat, dim3(n), dim3(256), 0, (hipStream_t)stream, curs, cw, refs, rs,
                       (const int2 *)pos, qps, feats);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvgpu_nn_train_step(const float *W, const float *feats, const int *tgt, const uint8_t *skip, int n,
                                    float wfp, float thr, float *grad, float *loss, int nblocks, void *stream) {
    static bool attr = false;
    if (!attr) {
        if (hipFuncSetAttribute((const void *)k_nn_train, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sizeof(NnLds)) != hipSuccess) return -1;
        attr = true;
    }
    hipLaunchKernelGGL(k_nn_train, dim3(nblocks), dim3(256), sizeof(NnLds), (hipStream_t)stream, W, feats, tgt, skip, n,
                       wfp, thr, grad, loss);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int fdvgpu_nn_adam(float *W, float *grad, float *m1, float *m2, int n, float lr, int step, int batch,
                              void *stream) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_nn_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, W, grad, m1, m2, n, lr,
                       b1, b2, 1e-8f, 1.f / (float)batch, c1, c2);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

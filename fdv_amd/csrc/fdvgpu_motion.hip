// fdvgpu_motion.hip -- motion search of one frame on caller-owned device
// buffers, outside any context: the block-parallel motion table of every MB
// (the exhaustive +-16 search, or the learned proposer's candidates with the
// exhaustive search below its confidence threshold).
//
// The reference searches one block per call (fdv_me_search, fdv.h:795-797,
// body 3375-3466); these entry points search every MB of a frame in one pass
// with the kernels the context-based frame step uses (k_nn_infer, then
// k_me_exhaustive).  Their scratch is the frame descriptor and, in the learned
// search, the per-MB proposals; fdvgpu_motion_search allocates it for the call
// and frees it once the launches that use it are done.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include "fdvgpu_common.h"
#include "../../include/fdvgpu.h"

extern "C" int fdvg_launch_metab(const fdvg_seq *, int, int, void *, const void *, float, float, hipStream_t);
extern "C" int fdvg_launch_nn_infer(const fdvg_seq *, int, int, const float *, fdvg_nnout *, int, hipStream_t);
extern "C" int fdvg_nn_setup(void);
extern "C" void fdvg_seterr(const char *fmt, ...);

#define MS_SKIP_OFF 2.0f                 // no SKIP hint: the table's flags stay 0

#define MS_DESC_BYTES 256
static_assert(sizeof(fdvg_seq) <= MS_DESC_BYTES, "frame descriptor");

// The descriptor is written by a one-lane kernel from its by-value argument,
// so no host buffer has to outlive the call.
__global__ void k_put_seq(fdvg_seq *dst, fdvg_seq v) { *dst = v; }

extern "C" int fdvg_motion_check(const uint8_t *cur_y, const uint8_t *ref0_y, const uint8_t *ref1_y, int ref_stride,
                                 int w, int h, int navail, int qp, const void *table) {
    if (!cur_y || !ref0_y || !table || w <= 0 || h <= 0 || (w & 15) || (h & 15) || w > 8192 || h > 8192 ||
        navail < 1 || navail > 2 || (navail == 2 && !ref1_y) || qp < 0 || qp > 51 ||
        ref_stride < w + 2 * FDVG_BORDER || (ref_stride & 3)) {
        fdvg_seterr("fdvgpu_motion_search: bad arguments");
        return -1;
    }
    return 0;
}

static fdvg_seq ms_desc(const uint8_t *cur_y, const uint8_t *ref0_y, const uint8_t *ref1_y, int ref_stride, int w,
                        int h, int navail, int qp) {
    fdvg_seq q;
    memset(&q, 0, sizeof q);
    q.cy = cur_y;
    q.ry[0] = ref0_y;
    q.ry[1] = navail == 2 ? ref1_y : ref0_y;
    q.w = w; q.h = h; q.ys = ref_stride;
    q.navail = navail; q.qp = qp;
    q.mb0 = 0;
    return q;
}

// The launches of one search: the descriptor, the proposer (learned), the table.
static int ms_launch(fdvg_seq *desc_d, fdvg_nnout *nn_d, const fdvg_seq &q, const float *weights, float conf,
                     fdvgpu_motion_entry *table, hipStream_t st) {
    const int mw = q.w >> 4, mh = q.h >> 4;
    hipLaunchKernelGGL(k_put_seq, dim3(1), dim3(1), 0, st, desc_d, q);
    if (hipGetLastError() != hipSuccess) { fdvg_seterr("fdvgpu_motion_search: descriptor launch failed"); return -1; }
    if (weights && fdvg_launch_nn_infer(desc_d, 1, mw * mh, weights, nn_d, 1, st) != 0) {
        fdvg_seterr("fdvgpu_motion_search: proposer launch failed");
        return -1;
    }
    if (fdvg_launch_metab(desc_d, 1, mh * ((mw + 7) >> 3) * 8, table, weights ? nn_d : nullptr, conf, MS_SKIP_OFF,
                          st) != 0) {
        fdvg_seterr("fdvgpu_motion_search: table launch failed");
        return -1;
    }
    return 0;
}

extern "C" int fdvgpu_motion_search(const uint8_t *cur_y, const uint8_t *ref0_y, const uint8_t *ref1_y,
                                    int ref_stride, int w, int h, int navail, int qp, const float *weights,
                                    float conf, fdvgpu_motion_entry *table, void *stream) {
    if (fdvg_motion_check(cur_y, ref0_y, ref1_y, ref_stride, w, h, navail, qp, table) != 0) return -1;
    if (weights && fdvg_nn_setup() != 0) { fdvg_seterr("fdvgpu_motion_search: proposer setup failed"); return -1; }
    const hipStream_t st = (hipStream_t)stream;
    const size_t nmb = (size_t)(w >> 4) * (h >> 4);
    uint8_t *buf = nullptr;
    if (hipMalloc(&buf, MS_DESC_BYTES + nmb * sizeof(fdvg_nnout)) != hipSuccess) {
        fdvg_seterr("fdvgpu_motion_search: device allocation failed");
        return -1;
    }
    const fdvg_seq q = ms_desc(cur_y, ref0_y, ref1_y, ref_stride, w, h, navail, qp);
    int rc = ms_launch((fdvg_seq *)buf, (fdvg_nnout *)(buf + MS_DESC_BYTES), q, weights, conf, table, st);
    // the scratch is freed only once the launches that use it are done
    if (hipStreamSynchronize(st) != hipSuccess && rc == 0) {
        fdvg_seterr("fdvgpu_motion_search: %s", hipGetErrorString(hipGetLastError()));
        rc = -1;
    }
    (void)hipFree(buf);
    return rc;
}

// The launches of one search into scratch the caller provides (the descriptor,
// MS_DESC_BYTES, then one proposal per MB).
extern "C" int fdvg_motion_launch(uint8_t *scratch_d, const uint8_t *cur_y, const uint8_t *ref0_y,
                                  const uint8_t *ref1_y, int ref_stride, int w, int h, int navail, int qp,
                                  const float *weights, float conf, fdvgpu_motion_entry *table, hipStream_t st) {
    if (weights && fdvg_nn_setup() != 0) { fdvg_seterr("fdvgpu_motion_search: proposer setup failed"); return -1; }
    const fdvg_seq q = ms_desc(cur_y, ref0_y, ref1_y, ref_stride, w, h, navail, qp);
    return ms_launch((fdvg_seq *)scratch_d, (fdvg_nnout *)(scratch_d + MS_DESC_BYTES), q, weights, conf, table, st);
}

// fdvgpu_noalloc.hip -- fdvgpu_motion_search without any allocation, for
// callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid it).
//
// The scratch of a search (the frame descriptor, then one proposer output per
// MB) lives in a fixed device array, `scratch`, sized for FDVGPU_NOALLOC_MAX_MB
// MBs; each device holds its own copy until the process ends.  Every call on a
// device shares that copy, so calls that overlap on one device race.  Nothing
// in it has to start at zero: each call writes the descriptor and, with
// weights, every MB's proposal before they are read.
#include <hip/hip_runtime.h>
#include <mutex>
#include "fdvgpu_common.h"
#include "../../include/fdvgpu.h"

extern "C" void fdvg_seterr(const char *fmt, ...);
extern "C" int fdvg_motion_check(const uint8_t *cur_y, const uint8_t *ref0_y, const uint8_t *ref1_y, int ref_stride,
                                 int w, int h, int navail, int qp, const void *table);
extern "C" int fdvg_motion_launch(uint8_t *scratch_d, const uint8_t *cur_y, const uint8_t *ref0_y,
                                  const uint8_t *ref1_y, int ref_stride, int w, int h, int navail, int qp,
                                  const float *weights, float conf, fdvgpu_motion_entry *table, hipStream_t st);

#define NA_DESC_BYTES 256               // the descriptor's slot (fdvgpu_motion.hip lays it out the same way)
static_assert(sizeof(fdvg_seq) <= NA_DESC_BYTES, "frame descriptor");
static __device__ __align__(256) uint8_t scratch[NA_DESC_BYTES + (size_t)FDVGPU_NOALLOC_MAX_MB * sizeof(fdvg_nnout)];

// Per device, the address of that device's `scratch`, looked up once with the
// device current.  In host code the array's name is only a stand-in, never a
// device address: kernels and copies get this pointer.
#define NA_MAXDEV 64
static std::mutex na_mu;
static uint8_t *na_ptr[NA_MAXDEV];

static uint8_t *na_scratch(void) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAXDEV) {
        fdvg_seterr("fdvgpu_motion_search_noalloc: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fdvg_seterr("fdvgpu_motion_search_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
            return nullptr;
        }
        na_ptr[dev] = (uint8_t *)p;
    }
    return na_ptr[dev];
}

extern "C" int fdvgpu_motion_search_noalloc(const uint8_t *cur_y, const uint8_t *ref0_y, const uint8_t *ref1_y,
                                            int ref_stride, int w, int h, int navail, int qp, const float *weights,
                                            float conf, fdvgpu_motion_entry *table, void *stream) {
    if (fdvg_motion_check(cur_y, ref0_y, ref1_y, ref_stride, w, h, navail, qp, table) != 0) return -1;
    if ((size_t)(w >> 4) * (h >> 4) > FDVGPU_NOALLOC_MAX_MB) {
        fdvg_seterr("fdvgpu_motion_search_noalloc: %dx%d has %d MBs, more than the %d its scratch holds", w, h,
                    (w >> 4) * (h >> 4), FDVGPU_NOALLOC_MAX_MB);
        return -1;
    }
    uint8_t *p = na_scratch();
    if (!p) return -1;
    return fdvg_motion_launch(p, cur_y, ref0_y, ref1_y, ref_stride, w, h, navail, qp, weights, conf, table,
                              (hipStream_t)stream);
}

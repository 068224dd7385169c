/* fdvgpu.h -- C ABI of the MI355X (gfx950) encode-analysis library.
 *
 * The reference codec (yhbsh/fdv, include/fdv.h) codes a P-frame in one call,
 *
 *     static size_t pframe_encode(cy, cu, cv, w, h, y0, bh, refs, navail, qp,
 *                                 out, cap, ry, ru, rv, force_skip, tc, lscale)
 *                                                          -- fdv.h:5764-6421
 *
 * whose MB loop (5808-6162) and band deblock (6168-6170) are the hot path; the
 * rest is host entropy coding (6172-6421).  This library replaces the hot path:
 * a context owns device memory for `nseq` independent sequences (a video, or one
 * band of a tile-parallel video -- fdv.h:7092-7138), and one fdvgpu_pframe call
 * analyses one frame of every sequence in a single batched launch sequence.  It
 * returns, per sequence, the 16 coded symbol streams that pframe_encode hands
 * to its entropy stage (fdv.h:6196-6207), the deblocked reconstruction, and the
 * per-MB decisions.  The host entropy packer (fdv_amd host library) turns the
 * streams into the reference's P-frame payload bytes.
 *
 * Conventions follow fdv.h: frames are tight I420 (w*h Y, then (w/2)*(h/2) U and
 * V), w and h are multiples of 16, qp is 0..51.  Functions return 0 (or a
 * pointer) on success and -1 (or NULL) on error -- never abort; the reason is
 * available from fdvgpu_last_error().  A context is not thread-safe; use one per
 * thread.
 */
#ifndef FDVGPU_H
#define FDVGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdvgpu fdvgpu;

/* Motion-search flavours for fdvgpu_pframe. */
enum {
    FDVGPU_ME_DIAMOND = 0,     /* the reference's EPZS diamond (fdv.h:3375): bit-exact   */
    FDVGPU_ME_EXHAUSTIVE = 1,  /* full +-16 pel integer search + the same sub-pel step  */
    FDVGPU_ME_LEARNED = 2,     /* proposer candidates, exhaustive fallback below 0.7,
                                  block-parallel decisions                             */
    FDVGPU_ME_PARALLEL = 3     /* exhaustive tables, block-parallel decisions           */
};

/* One MB decision (given-decisions input and fdvgpu_decisions output).
 * mode: 0 SKIP, 1 INTER16, 2 INTRA, 3 INTER8.  mv[] in quarter pels: [0] for
 * SKIP and INTER16, all four quadrants (raster) for INTER8.  merge_idx: which
 * merge candidate a SKIP inherits (fdv.h:5711).  sub: intra sub-mode 0..3. */
typedef struct {
    uint8_t mode, ref, merge_idx, sub;
    int16_t mv[4][2];
} fdvgpu_mbinfo;

/* Open a context for nseq sequences of the given sizes, with nsrc device slots
 * per sequence for source frames.  Replaces the per-call allocations of
 * pframe_encode (fdv.h:5782-5797) and the reference pool of fdv_video_encode
 * (fdv.h:6749-6754). */
fdvgpu *fdvgpu_open(int nseq, const int *w, const int *h, int nsrc);
void    fdvgpu_close(fdvgpu *g);
const char *fdvgpu_last_error(void);

/* Upload a tight I420 source frame into device slot `slot` of sequence `seq`. */
int fdvgpu_put_source(fdvgpu *g, int seq, int slot, const uint8_t *i420);

/* Install a reference frame (tight I420): which = 0 the nearer ("near"), 1 the
 * one before it ("far").  The equivalent of load_ref (fdv.h:5183) after a key
 * frame; P-frames rotate the pool themselves. */
int fdvgpu_put_ref(fdvgpu *g, int seq, int which, const uint8_t *i420);

/* Per-sequence decisions to use instead of search for the next fdvgpu_pframe
 * (nmb records in raster order), or NULL to search again. */
int fdvgpu_set_given(fdvgpu *g, int seq, const fdvgpu_mbinfo *dec);

/* Analyse + reconstruct one P-frame of every sequence (the MB loop and band
 * deblock of pframe_encode, fdv.h:5808-6170).  Per sequence: source slot, qp,
 * lambda scale (fdv.h:5804; 1.0 for fixed-QP coding), navail = references
 * available (1 or 2, fdv.h:6791), force_skip (a held frame, fdv.h:5820).  The
 * new reconstruction becomes "near" and the old near becomes "far"
 * (fdv.h:6842-6846).  Asynchronous: the results are read by the calls below. */
int fdvgpu_pframe(fdvgpu *g, const int *src_slot, const int *qp, const double *lscale,
                  const int *navail, const int *force_skip, int me_mode);

/* Wait for the last fdvgpu_pframe and fetch its symbol streams to the host.
 * Then fdvgpu_streams returns sequence seq's 16 streams: pointers into a
 * host buffer owned by the context (valid until the next fdvgpu_pframe) and
 * their lengths, in the order of fdv.h:6196-6207 (9 structure models, then
 * n l n8 l8 fl nc lc). */
int fdvgpu_fetch(fdvgpu *g);
int fdvgpu_streams(fdvgpu *g, int seq, const uint8_t **sv, size_t *sn);

/* Undo the reference rotation of the last fdvgpu_pframe, so that the same
 * frame can be coded again from the same references -- the requantising retry
 * of the streaming encoder (fdv.h:8303-8334), which calls pframe_encode twice
 * on one frame.  Valid once, directly after fdvgpu_pframe (fetching its results
 * first is allowed); -1 otherwise. */
int fdvgpu_rewind(fdvgpu *g);

/* The deblocked reconstruction of the last frame (tight I420, host memory). */
int fdvgpu_recon(fdvgpu *g, int seq, uint8_t *i420);

/* The per-MB decisions of the last frame (nmb records, raster order). */
int fdvgpu_decisions(fdvgpu *g, int seq, fdvgpu_mbinfo *out);

/* Wall-clock milliseconds of the device work of the last fdvgpu_pframe, from
 * HIP events on the context's stream: [0] motion tables (exhaustive / learned
 * modes), [1] the MB wavefront, [2] deblock + borders, [3] serialisation. */
int fdvgpu_timings(fdvgpu *g, float ms[4]);

#ifdef __cplusplus
}
#endif
#endif

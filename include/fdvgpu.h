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

/* Pipelined frame steps, for fixed-QP sequences whose next frame does not wait
 * on the previous frame's size (fdv_video_encode, fdv.h:6741-6854).  No
 * counterpart in the reference.  fdvgpu_submit uploads one frame of every
 * sequence (frames[s], tight I420; copied before it returns), queues its
 * analysis behind any earlier step, and stages its symbol streams in host
 * buffer `slot` (0 or 1); it returns without waiting for the device.
 * fdvgpu_collect waits for that slot's step only (later steps may still run)
 * and checks its flags; fdvgpu_collected_streams then returns sequence seq's
 * 16 streams, valid until the slot is submitted again.  A slot must be
 * collected before it is submitted again; with two slots, frame k + 1 is
 * analysed while the host packs frame k.  The reconstruction stays on the
 * device as the next reference; rewinding is not available after a submit. */
int fdvgpu_submit(fdvgpu *g, int slot, const uint8_t *const *frames, const int *qp, const double *lscale,
                  const int *navail, const int *force_skip, int me_mode);
int fdvgpu_collect(fdvgpu *g, int slot);
int fdvgpu_collected_streams(fdvgpu *g, int slot, int seq, const uint8_t **sv, size_t *sn);
/* The same pipelining for frames already in the context's source slots
 * (fdvgpu_put_source): src_slot[s] as in fdvgpu_pframe.  The step's
 * serialisation and the staging of its streams run on a stream of their own,
 * beside the next step's motion tables (whose decisions wait for them), so
 * that consecutive steps overlap on the device.  fdvgpu_collect /
 * fdvgpu_collected_streams as above; fdvgpu_slot_timings gives the step's
 * stage times (as fdvgpu_timings) once it is collected.  No counterpart in the
 * reference. */
int fdvgpu_submit_slots(fdvgpu *g, int slot, const int *src_slot, const int *qp, const double *lscale,
                        const int *navail, const int *force_skip, int me_mode);
int fdvgpu_slot_timings(fdvgpu *g, int slot, float ms[4]);

/* Undo the reference rotation of the last fdvgpu_pframe, so that the same
 * frame can be coded again from the same references -- the requantising retry
 * of the streaming encoder (fdv.h:8303-8334), which calls pframe_encode twice
 * on one frame.  Valid once, directly after fdvgpu_pframe (fetching its results
 * first is allowed); -1 otherwise. */
int fdvgpu_rewind(fdvgpu *g);

/* fdvgpu_rewind, and the next fdvgpu_pframe -- which must code the same source
 * frames in the same table mode (exhaustive, parallel or learned) -- reuses the
 * motion tables of the frame step it undoes instead of searching again: the
 * requantising retry at another QP then redoes only the decisions, coding,
 * deblock and serialisation.  The exhaustive tables do not depend on the QP,
 * so the retry's result equals a fresh step's; the learned mode keeps the
 * first attempt's proposals.  In the diamond mode the search depends on the
 * decisions and runs again.  The reference's retry (fdv.h:8303-8334) calls
 * pframe_encode again from the start. */
int fdvgpu_rewind_keep_motion(fdvgpu *g);

/* The deblocked reconstruction of the last frame (tight I420, host memory). */
int fdvgpu_recon(fdvgpu *g, int seq, uint8_t *i420);

/* The per-MB decisions of the last frame (nmb records, raster order). */
int fdvgpu_decisions(fdvgpu *g, int seq, fdvgpu_mbinfo *out);

/* ---- decoding.  One MB of a parsed P-frame: its decision and its quantised
 * levels, the form in which the symbol streams carry them (pframe_encode's
 * emission, fdv.h:6036-6160) and the reference's decoder reads them back
 * (pframe_decode, fdv.h:6492-6737).
 *   mode 0 SKIP, 1 INTER16, 2 INTRA, 3 INTER8; ref: the reference a SKIP or
 *   INTER16 MB predicts from; sub: the INTRA sub-mode (0..3);
 *   mv[k]: quarter-pel vectors (SKIP / INTER16 [0], INTER8 all four quadrants,
 *   raster); merge_idx, ncand, mvd: as coded (not read by the decoder);
 *   lflag[r]: luma 8x8 region r (raster): 0 nothing coded, 1 four 4x4
 *   blocks, 2 one 8x8 block; levels at lev[64 r ..]: four raster 4x4 blocks
 *   of 16 levels in raster order, or 64 levels of the 8x8 in raster order;
 *   ckind: the chroma coding of the mode: 0 none (SKIP), 1 an any-flag per
 *   plane then four 4x4 blocks (INTER16), 2 a region flag per plane as lflag
 *   (INTRA), 3 four 4x4 blocks per plane without a flag (INTER8); cflag[p]
 *   the flag of plane p (U, V); levels at clev[64 p ..] in lev's layout. */
typedef struct {
    uint8_t mode, ref, merge_idx, ncand;
    uint8_t sub, ckind, pad0, pad1;
    uint8_t lflag[4];
    uint8_t cflag[2];
    uint8_t pad2[2];
    int16_t mvd[4][2];
    int16_t mv[4][2];
    int16_t lev[256];
    int16_t clev[128];
} fdvgpu_mbcode;

/* Reconstruct one P-frame of every sequence on the device from its MB codes
 * (codes[s]: nmb records in raster order, host memory): prediction from the
 * near / far references (navail[s] of them) or intra neighbours, dequantised
 * inverse transforms, reconstruction and the band deblock at qp[s] -- the
 * device half of pframe_decode (fdv.h:6492-6737).  Vectors must pass the
 * reference decoder's bounds check (mv_in_bounds, fdv.h:6480).  The decoded
 * frame becomes the near reference, as after fdvgpu_pframe; read it with
 * fdvgpu_recon.  Synchronous. */
int fdvgpu_decode(fdvgpu *g, const fdvgpu_mbcode *const *codes, const int *qp, const int *navail);

/* ---- launch machinery.  Each frame step's device work after its descriptor
 * copies is a fixed launch sequence per (search mode, decision structure); from
 * its second use it is captured into a hipGraph and replayed (on by default).
 * Launch shapes (serialiser wavefronts per workgroup, deblock workgroup size,
 * proposer workgroups per CU, decision coding workers per CU) are chosen by fdvgpu_autotune, which times each
 * candidate on a real frame step (rewinding the reference pool after each) and
 * keeps the fastest.  None of this has a counterpart in the reference. */
int fdvgpu_set_graphs(fdvgpu *g, int on);

/* The structure of the block-parallel decision stage, all with the same
 * results: 0, the candidate costs, the decision replays and the coding of
 * their decisions as separate launches; 1, the costs, the replays (one
 * workgroup a MB row, consuming the cost words as they appear), then the
 * coding workers, on one stream; 2, the replays on a second stream beside the
 * costs and, after the costs, the coding workers beside the replays, each MB
 * coded as soon as it is decided (an intra MB's chroma once its neighbours are
 * coded); 3, the replays beside the costs as in 2, then as many coding workers
 * as in 1 once the replays are done; 4, as 2, and once the replays are done as
 * many more workers as in 3 join those beside them.  2 and 4 are available
 * where every replay still finds a CU beside the workers (at most 7/8 of the
 * CUs' worth of MB rows), and 4 is then the default, 3 otherwise;
 * fdvgpu_autotune keeps the fastest of 2, 3 and 4 unless the caller set the
 * structure here.  2 and 4 return -1 where unavailable.  No counterpart in the
 * reference. */
int fdvgpu_set_overlap(fdvgpu *g, int mode);
/* The decision structure in use (0 to 4), or -1 for a bad context. */
int fdvgpu_get_overlap(fdvgpu *g);
/* shapes: serialiser wavefronts per workgroup {1..8}, deblock workgroup size
 * {64, 128, 256}, proposer workgroups per CU {1, 2}, coding workers of the
 * fused decision stage per CU {1, 2}. */
int fdvgpu_get_tune(fdvgpu *g, int shapes[4]);
int fdvgpu_set_tune(fdvgpu *g, const int shapes[4]);
int fdvgpu_autotune(fdvgpu *g, const int *src_slot, const int *qp, const double *lscale,
                    const int *navail, int me_mode, int reps);

/* ---- the learned motion proposer (FDVGPU_ME_LEARNED).  It has no counterpart in
 * the reference: it chooses which whole-pel vectors the block-parallel search
 * evaluates, and which MBs go straight to SKIP; the codec arithmetic after it
 * is the reference's. */

/* One MB's proposal: 8 whole-pel vectors on the nearest reference (best first),
 * the 4 lowest-cost positions of the proposer's half-res correlation map
 * (whole pels; ties to the smaller |dx| + |dy|, then the lower position), the
 * 2 lowest of each 8x8 quadrant's own map (same ties), the 8 vectors' total
 * probability, and the SKIP probability. */
typedef struct {
    int8_t cx[8], cy[8];
    int8_t kx[4], ky[4];
    int8_t qx[4][2], qy[4][2];
    float conf, pskip;
} fdvgpu_proposal;

/* One MB's motion table entry (quarter-pel vectors): 16x16 per reference, the
 * four 8x8 quadrants on the nearest, their SADs; flags bit 0 = the proposer
 * expects SKIP, so the decision weighs SKIP against INTER16 only. */
typedef struct {
    int16_t mv16[2][2];
    int32_t sad16[2];
    int16_t mv8[4][2];
    int32_t sad8[4];
    int32_t flags;
} fdvgpu_motion_entry;

/* Number of proposer parameters (fp32), and installing them for the learned mode. */
int fdvgpu_nn_nparam(void);
int fdvgpu_set_proposer(fdvgpu *g, const float *weights, int n);
/* Proposer thresholds: an MB whose top-8 probability is at least `conf` (0.7 by
 * default) gets the candidate search, below it the exhaustive search; an MB
 * whose SKIP probability is at least `skip` is weighed SKIP against INTER16 only
 * (off by default: 2.0).  A value above 1 disables that use.  No counterpart in
 * the reference (the proposer is this library's own). */
int fdvgpu_set_proposer_thresholds(fdvgpu *g, float conf, float skip);

/* The proposals and the motion table of the last frame (nmb entries each). */
int fdvgpu_proposals(fdvgpu *g, int seq, fdvgpu_proposal *out);
int fdvgpu_motion(fdvgpu *g, int seq, fdvgpu_motion_entry *out);

/* Motion search of one frame on caller-owned device buffers, outside any
 * context: the motion table of every MB in one pass.  It replaces a loop of
 * fdv_me_search calls (fdv.h:795-797, body 3375-3466) over a frame's blocks
 * with the block-parallel search of this library's table modes:
 *   - weights NULL: the exhaustive +-16 pel search of every 16x16 block on each
 *     reference and of every 8x8 quadrant on the first, then the reference's
 *     sub-pel cross (FDVGPU_ME_EXHAUSTIVE's table);
 *   - weights (fdvgpu_nn_nparam() floats, device memory): the learned
 *     proposer's candidates where its confidence is at least `conf`, the
 *     exhaustive search elsewhere (FDVGPU_ME_LEARNED's table; flags stay 0).
 * cur_y: tight luma, stride w.  ref0_y / ref1_y: luma planes in the reference's
 * padded layout (fdv.h:1548: pointer at sample (0,0), 32 replicated samples on
 * every side, stride ref_stride >= w + 64, a multiple of 4); ref1_y is read
 * only when navail is 2.  qp: the frame's quantiser (the proposer's input
 * scale).  table: nmb entries of device memory, raster order.  Launched on
 * `stream` (a hipStream_t, or NULL); returns once the search is done.  Its
 * scratch (the frame descriptor and, with weights, the proposals) is
 * allocated for the call. */
int fdvgpu_motion_search(const uint8_t *cur_y, const uint8_t *ref0_y, const uint8_t *ref1_y, int ref_stride,
                         int w, int h, int navail, int qp, const float *weights, float conf,
                         fdvgpu_motion_entry *table, void *stream);

/* fdvgpu_motion_search without any allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * arguments and result; frames of at most FDVGPU_NOALLOC_MAX_MB macroblocks
 * (larger ones fail with an error).  Its scratch is a fixed device array that
 * each device holds for the life of the process and that every call on that
 * device shares: calls that overlap on one device race.  Asynchronous on
 * `stream`: the table is ready once the stream reaches this point. */
#define FDVGPU_NOALLOC_MAX_MB 34816      /* 4096 x 2176 */
int fdvgpu_motion_search_noalloc(const uint8_t *cur_y, const uint8_t *ref0_y, const uint8_t *ref1_y, int ref_stride,
                                 int w, int h, int navail, int qp, const float *weights, float conf,
                                 fdvgpu_motion_entry *table, void *stream);

/* Device pointers of a sequence's source slot luma (tight) and of the nearest
 * reference's padded luma (at its sample (0,0)), for feature extraction. */
int fdvgpu_planes(fdvgpu *g, int seq, int src_slot, const uint8_t **cur_y, int *cur_stride,
                  const uint8_t **ref_y, int *ref_stride);

/* Training on the GPU, on device buffers (stream: a hipStream_t, or NULL):
 * - fdvgpu_nn_features: the proposer input (289 x 2 floats) of n MBs, MB j at
 *   luma position pos[2j], pos[2j+1] of frame curs[j] (stride cw[j]) against the
 *   padded reference luma refs[j] (pointer to its sample (0,0), stride rs[j]),
 *   coded at QP qps[j];
 * - fdvgpu_nn_train_step: forward and backward over n tuples (features, vector
 *   target as a position 0..288 or -1, SKIP label), the SKIP loss weighing false
 *   positives by wfp, adding the summed gradient into grad and into loss[8]:
 *   vector loss, SKIP loss, top-1 hits, top-8 hits, tuples with SKIP
 *   probability >= thr and how many of them are SKIP, tuples with top-8
 *   probability >= thr and how many of them have the target in the top 8;
 * - fdvgpu_nn_adam: one Adam step (beta 0.9 / 0.999) with grad / batch, then
 *   grad cleared; step counts from 1. */
int fdvgpu_nn_features(const uint8_t *const *curs, const int *cw, const uint8_t *const *refs, const int *rs,
                       const int *pos, const int *qps, int n, float *feats, void *stream);
int fdvgpu_nn_train_step(const float *w, const float *feats, const int *tgt, const uint8_t *skip, int n,
                         float wfp, float thr, float *grad, float *loss, int nblocks, void *stream);
int fdvgpu_nn_adam(float *w, float *grad, float *m1, float *m2, int n, float lr, int step, int batch,
                   void *stream);

/* Wall-clock milliseconds of the device work of the last fdvgpu_pframe, from
 * HIP events on the context's stream: [0] motion tables (exhaustive / learned
 * modes), [1] the MB wavefront, [2] deblock + borders, [3] serialisation. */
int fdvgpu_timings(fdvgpu *g, float ms[4]);

/* ---- key frames on the device (fdv_image_encode, fdv.h:4845-4890)
 *
 * The rate-distortion decisions of the intra plane codec for n planes in one
 * device pass: the quadtree of 64x64 units, the leaves' modes and levels, the
 * reconstruction, deblocked.  in[i]: a plane of w x h samples (multiples of 4)
 * with row stride `stride`, coded at qp.  out[i]: host buffers the call fills:
 *   rec   w*h           the deblocked reconstruction;
 *   lev   w*h int16     every surviving transform block's levels in raster
 *                       order at the block's own position;
 *   rflag ceil(w/8)*ceil(h/8)   per 8x8 region of a leaf above 4x4: 0 empty,
 *                       1 four 4x4 transforms, 2 one 8x8;
 *   cmode (w/4)*(h/4)   per 4x4 cell the mode context value: a 4x4 leaf's
 *                       mode (0..8), or a larger leaf's mode mapped to 4x4
 *                       (vertical 1, horizontal 2, else 0);
 *   tree  fdvgpu_itree_bytes(w, h): for depth d = 0..3 (nodes of side 64 >> d
 *                       on a ceil(w/n) x ceil(h/n) grid) the split flags (0
 *                       leaf, 1 split, 2 split implied by the plane's edge),
 *                       then the leaf modes (DC, V, H, plane).
 * Returns 0, or -1 with fdvgpu_last_error set.  The process's device context
 * for key frames is created on the first call; calls are serialised. */
typedef struct { const uint8_t *src; int w, h, stride, qp; } fdvgpu_iplane_in;
typedef struct { uint8_t *rec; int16_t *lev; uint8_t *rflag, *cmode, *tree; } fdvgpu_itree_out;
size_t fdvgpu_itree_bytes(int w, int h);
int fdvgpu_image_batch(int n, const fdvgpu_iplane_in *in, fdvgpu_itree_out *out);
/* The host half: a decided plane (the layout above; rec is not read)
 * serialised into the codec's twelve streams and written as the image
 * container (fdv.h:4926-5030) into out (cap bytes).  Returns its length, or 0
 * with fdvamd_last_error set.  No device work. */
size_t fdvgpu_image_pack(const fdvgpu_itree_out *t, int w, int h, int qp, uint8_t *out, size_t cap);

#ifdef __cplusplus
}
#endif
#endif

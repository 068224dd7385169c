/* fdv_amd.h -- host C11 side of the MI355X encode-analysis path.
 *
 * The drop-in point is the reference's P-frame encoder call
 *
 *   fdv.h:6793   blob = pframe_encode(cy, cu, cv, w, h, 0, h, refs, navail, qp, tmp,
 *                                     tmp_cap, ry, ru, rv, 0, &tc, 1.0);
 *   fdv.h:8278   (the streaming encoder's call, same arguments + force_skip, lscale)
 *
 * fdvamd_pframe_encode takes the same arguments except the reference frames,
 * which the encoder keeps resident in device memory (the reference's near / far
 * pool, fdv.h:6749-6754, 6842-6846): after a key frame the caller installs the
 * key frame's reconstruction with fdvamd_set_ref, exactly where the reference
 * calls load_ref (fdv.h:6844).  The output bytes, the reconstruction and the
 * entropy state `tc` are identical to pframe_encode's.
 *
 * fdvamd_pframe_batch codes one frame of many sequences at once -- the bands of
 * a tile-parallel video (fdv_vtile_encode, fdv.h:7161) or independent streams
 * -- in a single batched device pass, then packs each sequence's payload on
 * host threads.
 *
 * The entropy coder is the reference's (static rANS with four table strategies
 * and the adaptive range coder, fdv.h:1747-2143, 6172-6421) restated in C11; it
 * is exposed as fdvamd_pack_pframe for callers that bring their own streams.
 * Errors: sizes are 0 and status codes -1, as in fdv.h.
 */
#ifndef FDV_AMD_H
#define FDV_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Entropy state carried between P-frames; the same layout as fdv_tabcache
 * (fdv.h:599-612), so a reference encoder's cache can be passed straight in. */
typedef struct {
    uint16_t f1[256];
    uint16_t fs[16][256];
    int      have;
    uint32_t hist[16][256];
    int      hn;
} fdvamd_tabcache;

/* Entropy-code a P-frame from its 16 symbol streams (fdv.h:6172-6421).
 * Returns the payload length, or 0 if it does not fit in cap. */
size_t fdvamd_pack_pframe(int qp, const uint8_t *const sv[16], const size_t sn[16],
                          fdvamd_tabcache *tc, uint8_t *out, size_t cap);

typedef struct fdvamd_encoder fdvamd_encoder;

/* An encoder for nseq sequences of the given sizes (multiples of 16).
 * me_mode: 0 the reference's diamond search (bit-exact with pframe_encode),
 * 1 exhaustive search, 2 learned proposer with exhaustive fallback and
 * block-parallel decisions, 3 exhaustive search with block-parallel decisions. */
fdvamd_encoder *fdvamd_encoder_open(int nseq, const int *w, const int *h, int me_mode);
void fdvamd_encoder_close(fdvamd_encoder *e);
const char *fdvamd_last_error(void);

/* Install a reference (tight I420): which = 0 near, 1 far (load_ref, fdv.h:5183). */
int fdvamd_set_ref(fdvamd_encoder *e, int seq, int which, const uint8_t *i420);

/* The same from three tight planes (w x h luma, w/2 x h/2 chroma), as the
 * reference's encoders hold their reconstruction (ry/ru/rv, fdv.h:6766-6770):
 * after an I-frame, install its reconstruction as the near reference. */
int fdvamd_set_ref_planes(fdvamd_encoder *e, int seq, int which,
                          const uint8_t *y, const uint8_t *u, const uint8_t *v);

/* pframe_encode for sequence 0 of the encoder (fdv.h:5764): returns the payload
 * length, or 0 on error.  ry/ru/rv (tight, may be NULL) receive the deblocked
 * reconstruction; tc (may be NULL) is the entropy state. */
size_t fdvamd_pframe_encode(fdvamd_encoder *e, const uint8_t *cy, const uint8_t *cu,
                            const uint8_t *cv, int w, int h, int navail, int qp,
                            uint8_t *out, size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv,
                            int force_skip, fdvamd_tabcache *tc, double lscale);

/* One frame of every sequence: frames[s] tight I420; out[s]/cap[s] payload
 * buffers; len[s] receives each payload length (0 on error); recon[s] (may be
 * NULL) the deblocked reconstruction; tc[s] (may be NULL) entropy states.
 * Returns 0, or -1 if any sequence failed. */
int fdvamd_pframe_batch(fdvamd_encoder *e, const uint8_t *const *frames, const int *navail,
                        const int *qp, const double *lscale, const int *force_skip,
                        fdvamd_tabcache *const *tc, uint8_t *const *out, const size_t *cap,
                        size_t *len, uint8_t *const *recon);

/* Pipelined P-frames for fixed-QP sequences (fdv_video_encode, fdv.h:6741-6854,
 * whose next frame does not wait on the previous frame's size).  No counterpart
 * in the reference.  fdvamd_pframe_submit queues one frame of every sequence
 * (frames copied before it returns) and returns without waiting for the device;
 * at most two frames are in flight.  fdvamd_pframe_collect returns the oldest
 * submitted frame's payloads (as fdvamd_pframe_batch, without the
 * reconstruction, which stays on the device as the next reference) while the
 * next frame is analysed.  Order: submit 1, submit 2, collect 1, submit 3,
 * collect 2, ...  fdvamd_pframe_batch, fdvamd_pframe_encode and fdvamd_rewind
 * fail while frames are in flight.  Both return 0, or -1. */
int fdvamd_pframe_submit(fdvamd_encoder *e, const uint8_t *const *frames, const int *navail, const int *qp,
                         const double *lscale, const int *force_skip);
int fdvamd_pframe_collect(fdvamd_encoder *e, fdvamd_tabcache *const *tc, uint8_t *const *out, const size_t *cap,
                          size_t *len);

/* Install the learned proposer's weights (fdvgpu_nn_nparam() floats) for
 * me_mode 2.  No counterpart in the reference. */
int fdvamd_set_proposer(fdvamd_encoder *e, const float *weights, int n);

/* The proposer's confidence and SKIP thresholds (fdvgpu_set_proposer_thresholds).
 * No counterpart in the reference. */
int fdvamd_set_proposer_thresholds(fdvamd_encoder *e, float conf, float skip);

/* Undo the last frame's reference rotation (all sequences) so the frame can be
 * coded again, e.g. at another QP: the requantising retry of the streaming
 * encoder (fdv.h:8303-8334).  The caller restores its entropy state as the
 * reference does (tc_before, fdv.h:8273).  Returns 0, or -1 if there is no
 * frame to undo. */
int fdvamd_rewind(fdvamd_encoder *e);

/* fdvamd_rewind for a retry of the same frames: the retry (fdvamd_pframe_batch
 * or fdvamd_pframe_encode with the same frames, any QP) reuses the undone
 * step's motion tables in the table modes (fdvgpu_rewind_keep_motion): the
 * "a requantising retry should reuse ME" of fdv.h:8303-8334.  Returns 0, or -1
 * if there is no frame to undo. */
int fdvamd_rewind_keep_motion(fdvamd_encoder *e);

/* ---- sequences, with the reference's signatures and containers ---------- */

/* ---- key frames on the device
 *
 * fdvamd_image_encode replaces fdv_image_encode (fdv.h:871, body 4845-5038),
 * same arguments and result: the intra-only plane codec of w x h samples
 * (multiples of 4) with row stride `stride` at qp; writes the container to
 * out (cap bytes) and the deblocked reconstruction (w x h, tight) to recon
 * when not NULL; returns the container's length, or 0.  The rate-distortion
 * quadtree runs on the device, the serialisation and entropy coding on the
 * host.  fdvamd_image_encode_batch codes n planes in one device pass (len[i]
 * the length of plane i; returns 0, or -1 with fdvamd_last_error set).
 * fdvamd_iframe_encode codes a whole key frame -- the three planes of a tight
 * I420 frame, each behind its u32 little-endian length, as iframe_encode
 * (fdv.h:5487-5539) -- with the contract of fdvamd_iframe_encode_fn; it is the
 * key-frame coder of the sequence and streaming encoders unless another is
 * registered with fdvamd_set_iframe_coder. */
size_t fdvamd_image_encode(const uint8_t *src, int w, int h, int stride, int qp, uint8_t *out, size_t cap,
                           uint8_t *recon);
int fdvamd_image_encode_batch(int n, const uint8_t *const *src, const int *w, const int *h, const int *stride,
                              const int *qp, uint8_t *const *out, const size_t *cap, uint8_t *const *recon,
                              size_t *len);
size_t fdvamd_iframe_encode(const uint8_t *y, const uint8_t *u, const uint8_t *v, int w, int h, int qp, uint8_t *out,
                            size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv);
/* fdvamd_image_decode replaces fdv_image_decode (fdv.h:5040-5152), same
 * arguments and result: one plane's container decoded into dst (row stride
 * `stride`, at least the plane's width), its size in *w, *h when not NULL;
 * returns 0, or -1 for input the reference decoder rejects.  On the host.
 * fdvamd_iframe_decode decodes a whole key frame (iframe_decode, fdv.h:5541)
 * with the contract of fdvamd_iframe_decode_fn; it is the key-frame decoder of
 * the sequence and streaming decoders unless another is registered. */
int fdvamd_image_decode(const uint8_t *in, size_t len, uint8_t *dst, int stride, int *w, int *h);
int fdvamd_iframe_decode(const uint8_t *in, size_t len, uint8_t *y, uint8_t *u, uint8_t *v, int w, int h);

/* A key-frame coder: codes the tight planes y (w x h), u, v (w/2 x h/2) at qp
 * into out (at most cap bytes) and leaves its reconstruction in ry, ru, rv;
 * returns the blob length, or 0 on failure -- the contract of the reference's
 * iframe_encode (fdv.h:5487) over a whole frame.  The decoder's counterpart
 * (iframe_decode, fdv.h:5541) returns 0, or -1 on a malformed blob.  Key
 * frames are coded on the device (fdvamd_iframe_encode) and decoded on the host
 * (fdvamd_iframe_decode) unless coders are registered here (process-wide;
 * either may be NULL to fall back to the built-in one).  No counterpart in the reference, whose sequence encoders call
 * their own. */
typedef size_t (*fdvamd_iframe_encode_fn)(const uint8_t *y, const uint8_t *u, const uint8_t *v, int w, int h, int qp,
                                          uint8_t *out, size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv);
typedef int (*fdvamd_iframe_decode_fn)(const uint8_t *in, size_t len, uint8_t *y, uint8_t *u, uint8_t *v, int w,
                                       int h);
int fdvamd_set_iframe_coder(fdvamd_iframe_encode_fn enc, fdvamd_iframe_decode_fn dec);

/* The search / decision mode of the sequence entry points below (process-wide,
 * as fdvamd_encoder_open's me_mode; default 0, the reference's diamond search:
 * their output is then byte-identical to the reference's).  No counterpart in
 * the reference. */
int fdvamd_set_sequence_mode(int me_mode);
/* The learned proposer's weights for mode 2 of the sequence entry points
 * (copied; NULL clears them).  No counterpart in the reference. */
int fdvamd_set_sequence_proposer(const float *weights, int n);

/* fdv_video_encode (fdv.h:903-904, body 6741-6854), same signature and output:
 * the 8-byte sequence header (frame count, width, height, qp, keyint clamped
 * to 0..255), then per frame a u32 length and the frame's blob; frame 0 and
 * every keyint-th frame are key frames (coded on the device, or by the
 * registered coder), the others
 * P-frames coded on the device against a two-frame reference pool held there.
 * Returns the stream length, or 0 on error (fdvamd_last_error() says why). */
size_t fdvamd_video_encode(const uint8_t *const *frames, int nframes, int w, int h, int qp, int keyint,
                           uint8_t *out, size_t cap);

/* fdv_vtile_encode (fdv.h:973-975, body 7161-7243), same signature and output:
 * horizontal bands of band_mbrows MB rows, each an independent
 * fdvamd_video_encode stream (with the reference's per-band capacity of two
 * band frames + 64 KiB), behind the 16-byte tiled header and per-band u32
 * lengths in band order.  Every band's P-frame is coded in one device pass;
 * nthreads bounds the host threads of the key frames. */
size_t fdvamd_vtile_encode(const uint8_t *const *frames, int nframes, int w, int h, int qp, int keyint,
                           int band_mbrows, int nthreads, uint8_t *out, size_t cap);

/* ---- decoding ----------------------------------------------------------- */

/* The inverse of fdvamd_pack_pframe: a P-frame payload's entropy coding undone
 * (pframe_decode's first half, fdv.h:6492-6625) into its 16 symbol streams,
 * stored one after another in *syms (malloc'd, the caller frees it) with their
 * lengths in sn[]; *qp is the frame's quantiser.  nmb (the frame's MB count)
 * bounds the symbol count as the reference does.  tc (may be NULL) is the
 * entropy state carried between P-frames, updated as the reference decoder
 * updates it.  Returns 0, or -1 for a payload the reference decoder rejects. */
int fdvamd_unpack_pframe(const uint8_t *in, size_t len, size_t nmb, fdvamd_tabcache *tc, int *qp, uint8_t **syms,
                         size_t sn[16]);

/* The MB syntax of a P-frame's streams (pframe_decode, fdv.h:6626-6729; the
 * residual syntax of 5280, 5426, 6429) parsed into one fdvgpu_mbcode per MB
 * (include/fdvgpu.h), raster order, for a w x h frame.  Returns 0, or -1 for
 * streams the reference decoder rejects (including vectors outside its
 * bounds check, fdv.h:6480). */
int fdvamd_parse_pframe(const uint8_t *syms, const size_t sn[16], int w, int h, void *codes);

typedef struct fdvamd_decoder fdvamd_decoder;

/* A P-frame decoder for nseq sequences of the given sizes: a two-frame
 * reference pool per sequence in device memory (fdv_video_decode's pool,
 * fdv.h:6870-6874).  fdvamd_decoder_set_ref installs a decoded key frame
 * (tight I420) as the near reference (load_ref + pool swap, 6921-6923). */
fdvamd_decoder *fdvamd_decoder_open(int nseq, const int *w, const int *h);
void fdvamd_decoder_close(fdvamd_decoder *d);
int fdvamd_decoder_set_ref(fdvamd_decoder *d, int seq, const uint8_t *i420);

/* One P-frame of every sequence: payload in[s] (len[s] bytes) against navail[s]
 * references, entropy state tc[s] (may be NULL); the decoded frame (tight
 * I420) into out[s] (may be NULL: it stays on the device as the next
 * reference).  pframe_decode (fdv.h:6492-6737) for all sequences: the payloads
 * are unpacked and parsed on host threads, the frames reconstructed and
 * deblocked on the device in one pass.  Returns 0, or -1 if any payload is
 * malformed. */
int fdvamd_pframe_decode_batch(fdvamd_decoder *d, const uint8_t *const *in, const size_t *len, const int *navail,
                               fdvamd_tabcache *const *tc, uint8_t *const *out);

/* fdv_video_decode (fdv.h:907-908, body 6857-6929) and fdv_vtile_decode
 * (980-981, body 7280-7378), same signatures and results: all nframes decoded
 * frames, consecutive tight I420, into out; key frames from the registered
 * decoder (fdvamd_set_iframe_coder), P-frames on the device, every band of a
 * tiled frame in one pass (the bands of a tiled stream must share their key
 * interval, as the reference's tiled encoder writes them).  Returns 0, or -1
 * for a stream the reference decoder rejects. */
int fdvamd_video_decode(const uint8_t *in, size_t len, uint8_t *out, int *nframes, int *w, int *h);
int fdvamd_vtile_decode(const uint8_t *in, size_t len, uint8_t *out, int *nframes, int *w, int *h, int nthreads);

/* ---- file container, streaming decode and encode, bitrate control ---------
 * The reference's public streaming surface with its signatures and semantics;
 * the P-frames of every band of a frame go through the device in one pass, key
 * frames through the device key-frame coder (or the registered one) and the
 * registered I-frame decoder. */

/* fdv_info (fdv.h:1076-1085), same layout. */
typedef struct {
    int kind;
    int version;
    int fps;
    int w, h, nframes;
    int qp, keyint;
    int nbands;
    const uint8_t *payload;
    size_t payload_len;
} fdvamd_info;
#define FDVAMD_HEAD 12
enum { FDVAMD_VIDEO = 0, FDVAMD_TILED = 1 };

/* fdv_wrap / fdv_read / fdv_decode (fdv.h:1090-1105, bodies 7448-7520): the
 * 12-byte "FDV1" head around a video or tiled stream. */
size_t fdvamd_wrap(uint8_t *out, size_t cap, int kind, int fps, const uint8_t *payload, size_t payload_len);
int fdvamd_read(const uint8_t *in, size_t len, fdvamd_info *info);
int fdvamd_decode(const fdvamd_info *info, uint8_t *out, int nthreads);

/* fdv_dec_* (fdv.h:1126-1142, bodies 7526-7795): a frame at a time, seekable;
 * the bands of a tiled file step together (they must share their key
 * interval, as the reference's tiled encoder writes them). */
typedef struct fdvamd_dec fdvamd_dec;
fdvamd_dec *fdvamd_dec_open(const fdvamd_info *info);
int fdvamd_dec_next(fdvamd_dec *d, uint8_t *out);
int fdvamd_dec_seek(fdvamd_dec *d, int frame);
int fdvamd_dec_pos(const fdvamd_dec *d);
size_t fdvamd_dec_frame_size(const fdvamd_dec *d);
void fdvamd_dec_close(fdvamd_dec *d);

/* fdv_enc_* (fdv.h:1146-1198, bodies 7798-8459): a frame at a time, with held
 * (all-SKIP) frames, average-bitrate control and its requantising retry (which
 * keeps the device's motion tables); the output equals fdv_enc_finish's. */
typedef struct fdvamd_enc fdvamd_enc;
fdvamd_enc *fdvamd_enc_open(int w, int h, int qp, int keyint);
fdvamd_enc *fdvamd_enc_open_tiled(int w, int h, int qp, int keyint, int band_mbrows, int nthreads);
int fdvamd_enc_frame(fdvamd_enc *e, const uint8_t *i420);
int fdvamd_enc_repeat(fdvamd_enc *e);
int fdvamd_enc_set_bitrate(fdvamd_enc *e, int bits_per_second, int fps, double bufsecs);
int fdvamd_enc_count(const fdvamd_enc *e);
size_t fdvamd_enc_bytes(const fdvamd_enc *e);
uint8_t *fdvamd_enc_finish(fdvamd_enc *e, int fps, size_t *len);
void fdvamd_enc_close(fdvamd_enc *e);

/* fdv_rc (fdv.h:1023-1054, bodies 7880-8040), same layout and arithmetic. */
typedef struct {
    int bitrate;
    int fps;
    double bufbits;
    int qpmin, qpmax, keyint;
    double qp_p;
    double ip_off;
    double last_target;
    double fill;
    int used_qp;
    double lscale;
} fdvamd_rc;
void fdvamd_rc_init(fdvamd_rc *rc, int bitrate, int fps, int w, int h, int keyint, double bufsecs, int qpmin,
                    int qpmax);
int fdvamd_rc_pick(fdvamd_rc *rc, int is_intra);
void fdvamd_rc_update(fdvamd_rc *rc, int is_intra, size_t bytes);
void fdvamd_rc_skip(fdvamd_rc *rc, size_t bytes);
void fdvamd_rc_adopt(fdvamd_rc *rc, int is_intra, double q);
void fdvamd_rc_split(const fdvamd_rc *rc, double q, int *qp, double *lscale);

#ifdef __cplusplus
}
#endif
#endif

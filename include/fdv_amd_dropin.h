/* fdv_amd_dropin.h -- the reference's sequence coders, file container,
 * streaming encoder and decoder on the MI355X path, for a caller that changes
 * nothing but one include.
 *
 * Include it after fdv.h in a source file whose calls to
 *   fdv_image_encode / fdv_image_decode  (fdv.h:871-876)
 *   fdv_video_encode / fdv_video_decode  (fdv.h:903-908)
 *   fdv_vtile_encode / fdv_vtile_decode  (fdv.h:973-981)
 *   fdv_wrap / fdv_read / fdv_decode     (fdv.h:1090-1105)
 *   fdv_dec_*                            (fdv.h:1126-1142)
 *   fdv_enc_*                            (fdv.h:1146-1198)
 *   fdv_rc_init / _pick / _update / _skip / _adopt / _split (fdv.h:1023-1054)
 * should go to this library: the calls after the include are renamed to the
 * fdvamd_ functions of the same signatures (include/fdv_amd.h), which write and
 * read the same containers; fdvamd_info, fdvamd_rc and the encoder / decoder
 * handles stand for fdv_info, fdv_rc, fdv_encoder and fdv_decoder.  Key frames
 * are coded by this library too (on the device) and decoded by it (on the
 * host).  A file that defines FDV_IMPLEMENTATION (where fdv.h's file-static
 * iframe_encode / iframe_decode, fdv.h:5487 / 5541, are visible) may call
 * fdvamd_dropin_register() to route key frames to the reference's own intra
 * coder instead.  Link with fdv_amd/libfdvgpu.so.
 */
#ifndef FDV_AMD_DROPIN_H
#define FDV_AMD_DROPIN_H
#include "fdv_amd.h"

#ifdef FDV_IMPLEMENTATION
static size_t fdvamd_dropin_ienc(const uint8_t *y, const uint8_t *u, const uint8_t *v, int w, int h, int qp,
                                 uint8_t *out, size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv) {
    return iframe_encode(y, u, v, w, h, 0, h, qp, out, cap, ry, ru, rv);
}
static int fdvamd_dropin_idec(const uint8_t *in, size_t len, uint8_t *y, uint8_t *u, uint8_t *v, int w, int h) {
    return iframe_decode(in, len, y, u, v, w, h, 0, h);
}
static inline int fdvamd_dropin_register(void) {
    return fdvamd_set_iframe_coder(fdvamd_dropin_ienc, fdvamd_dropin_idec);
}
#endif

#define fdv_image_encode fdvamd_image_encode
#define fdv_image_decode fdvamd_image_decode
#define fdv_video_encode fdvamd_video_encode
#define fdv_video_decode fdvamd_video_decode
#define fdv_vtile_encode fdvamd_vtile_encode
#define fdv_vtile_decode fdvamd_vtile_decode

#define fdv_wrap fdvamd_wrap
#define fdv_read(in, len, info) fdvamd_read((in), (len), (fdvamd_info *)(info))
#define fdv_decode(info, out, nthreads) fdvamd_decode((const fdvamd_info *)(info), (out), (nthreads))

#define fdv_dec_open(info) ((fdv_decoder *)fdvamd_dec_open((const fdvamd_info *)(info)))
#define fdv_dec_next(d, out) fdvamd_dec_next((fdvamd_dec *)(d), (out))
#define fdv_dec_seek(d, frame) fdvamd_dec_seek((fdvamd_dec *)(d), (frame))
#define fdv_dec_pos(d) fdvamd_dec_pos((const fdvamd_dec *)(d))
#define fdv_dec_frame_size(d) fdvamd_dec_frame_size((const fdvamd_dec *)(d))
#define fdv_dec_close(d) fdvamd_dec_close((fdvamd_dec *)(d))

#define fdv_enc_open(w, h, qp, keyint) ((fdv_encoder *)fdvamd_enc_open((w), (h), (qp), (keyint)))
#define fdv_enc_open_tiled(w, h, qp, keyint, band_mbrows, nthreads) \
    ((fdv_encoder *)fdvamd_enc_open_tiled((w), (h), (qp), (keyint), (band_mbrows), (nthreads)))
#define fdv_enc_frame(e, i420) fdvamd_enc_frame((fdvamd_enc *)(e), (i420))
#define fdv_enc_repeat(e) fdvamd_enc_repeat((fdvamd_enc *)(e))
#define fdv_enc_set_bitrate(e, bps, fps, bufsecs) fdvamd_enc_set_bitrate((fdvamd_enc *)(e), (bps), (fps), (bufsecs))
#define fdv_enc_count(e) fdvamd_enc_count((const fdvamd_enc *)(e))
#define fdv_enc_bytes(e) fdvamd_enc_bytes((const fdvamd_enc *)(e))
#define fdv_enc_finish(e, fps, len) fdvamd_enc_finish((fdvamd_enc *)(e), (fps), (len))
#define fdv_enc_close(e) fdvamd_enc_close((fdvamd_enc *)(e))

#define fdv_rc_init(rc, bitrate, fps, w, h, keyint, bufsecs, qpmin, qpmax) \
    fdvamd_rc_init((fdvamd_rc *)(rc), (bitrate), (fps), (w), (h), (keyint), (bufsecs), (qpmin), (qpmax))
#define fdv_rc_pick(rc, is_intra) fdvamd_rc_pick((fdvamd_rc *)(rc), (is_intra))
#define fdv_rc_update(rc, is_intra, bytes) fdvamd_rc_update((fdvamd_rc *)(rc), (is_intra), (bytes))
#define fdv_rc_skip(rc, bytes) fdvamd_rc_skip((fdvamd_rc *)(rc), (bytes))
#define fdv_rc_adopt(rc, is_intra, q) fdvamd_rc_adopt((fdvamd_rc *)(rc), (is_intra), (q))
#define fdv_rc_split(rc, q, qp, lscale) fdvamd_rc_split((const fdvamd_rc *)(rc), (q), (qp), (lscale))

#endif

/* fdv_amd_dropin.h -- the reference's sequence coders on the MI355X path, for a
 * caller that changes nothing but one include.
 *
 * Include it after fdv.h in a source file whose calls to
 *   fdv_video_encode / fdv_video_decode  (fdv.h:903-908)
 *   fdv_vtile_encode / fdv_vtile_decode  (fdv.h:973-981)
 * should go to this library: the calls after the include are renamed to the
 * fdvamd_ functions of the same signatures (include/fdv_amd.h), which write and
 * read the same containers.  Key frames stay with the reference's intra coder:
 * in the one file that defines FDV_IMPLEMENTATION (where fdv.h's file-static
 * iframe_encode / iframe_decode, fdv.h:5487 / 5541, are visible), call
 * fdvamd_dropin_register() once before the first encode or decode.  Link with
 * fdv_amd/libfdvgpu.so.
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

#define fdv_video_encode fdvamd_video_encode
#define fdv_video_decode fdvamd_video_decode
#define fdv_vtile_encode fdvamd_vtile_encode
#define fdv_vtile_decode fdvamd_vtile_decode

#endif

/*
 * ref_driver.c — drives the reference's own compiled CPU reconstruction
 * (vp9_reconstruct_frame, cvp9_frame_copy from src/decoder/) through the
 * decoder's reference-pool rule of src/decoder/vp9_decoder.c:209-231.
 * TEST INFRASTRUCTURE ONLY: built into oracle/_ref by `make ref`, used to
 * validate the C restatement and as the timed CPU baseline.
 */
#include <stdlib.h>
#include <string.h>
#include "compute_vp9/decoder.h"
#include "vp9_reconstruct.h"
#include "vp9_frame.h"

typedef struct { cvp9_frame_info_t refs[8]; } ref_decoder_t;

ref_decoder_t *ref_decoder_new(void) { return calloc(1, sizeof(ref_decoder_t)); }

void ref_decoder_free(ref_decoder_t *d)
{
    if (!d) return;
    for (int i = 0; i < 8; i++) cvp9_frame_free(&d->refs[i]);
    free(d);
}

/* Runs n frames; writes each frame's packed I420 to out (NULL: discard). */
int ref_decoder_run(ref_decoder_t *d, const vp9_parsed_frame_t *const *pfs, int n, uint8_t *out)
{
    int done = 0;
    for (int i = 0; i < n; i++) {
        cvp9_frame_info_t f;
        memset(&f, 0, sizeof(f));
        const cvp9_frame_info_t *refs[8];
        for (int k = 0; k < 8; k++) refs[k] = &d->refs[k];
        if (!vp9_reconstruct_frame(pfs[i], refs, &f)) break;
        f.pts = i;
        if (pfs[i]->hdr.frame_type == VP9_FRAME_KEY)
            for (int k = 0; k < 8; k++) cvp9_frame_copy(&d->refs[k], &f);
        else
            cvp9_frame_copy(&d->refs[0], &f);
        if (out) {
            const uint32_t cw = (f.width + 1) / 2, ch = (f.height + 1) / 2;
            memcpy(out, f.plane_y, (size_t)f.width * f.height);
            out += (size_t)f.width * f.height;
            memcpy(out, f.plane_u, (size_t)cw * ch);
            out += (size_t)cw * ch;
            memcpy(out, f.plane_v, (size_t)cw * ch);
            out += (size_t)cw * ch;
        }
        cvp9_frame_free(&f);
        done++;
    }
    return done;
}

/* One frame against caller-provided references, with no pool update. */
int ref_reconstruct(const vp9_parsed_frame_t *pf, const cvp9_frame_info_t *const refs[8],
                    uint8_t *out)
{
    cvp9_frame_info_t f;
    memset(&f, 0, sizeof(f));
    if (!vp9_reconstruct_frame(pf, (const cvp9_frame_info_t **)refs, &f)) return -1;
    const uint32_t cw = (f.width + 1) / 2, ch = (f.height + 1) / 2;
    memcpy(out, f.plane_y, (size_t)f.width * f.height);
    memcpy(out + (size_t)f.width * f.height, f.plane_u, (size_t)cw * ch);
    memcpy(out + (size_t)f.width * f.height + (size_t)cw * ch, f.plane_v, (size_t)cw * ch);
    cvp9_frame_free(&f);
    return 0;
}

/* Renders the reference's scene files (include/fdv_scene.h: fdv_scene_parse_file,
 * fdv_scene_resize, fdv_scene_render) to raw I420, for the rate-distortion study
 * of the search modes.  Content only: nothing here is on the encode path.
 *
 *     scene_render <file.scn> <width> <height> <frames> <out.yuv>
 *
 * Built by fdv_amd/learned/Makefile from the scene header where it lies. */
#define FDV_SCENE_IMPLEMENTATION
#include "fdv_scene.h"

#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: scene_render <file.scn> <width> <height> <frames> <out.yuv>\n");
        return 2;
    }
    char err[256] = "";
    fdv_scene *s = fdv_scene_parse_file(argv[1], err, sizeof err);
    if (!s) { fprintf(stderr, "scene %s: %s\n", argv[1], err); return 1; }
    const int w = atoi(argv[2]), h = atoi(argv[3]), nf = atoi(argv[4]);
    if (nf <= 0 || fdv_scene_resize(s, w, h) != 0) {
        fprintf(stderr, "bad size %dx%d or frame count %d\n", w, h, nf);
        fdv_scene_free(s);
        return 1;
    }
    const size_t fs = fdv_scene_frame_size(s);
    uint8_t *buf = malloc(fs);
    FILE *f = fopen(argv[5], "wb");
    if (!buf || !f) { fprintf(stderr, "cannot write %s\n", argv[5]); return 1; }
    for (int i = 0; i < nf; ++i) {
        fdv_scene_render(s, i, buf);
        if (fwrite(buf, 1, fs, f) != fs) { fprintf(stderr, "short write\n"); return 1; }
    }
    fclose(f);
    free(buf);
    fdv_scene_free(s);
    return 0;
}

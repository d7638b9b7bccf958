/* stb_bmp_probe.c -- BMP ground truth from the stb_image_write v1.14 that rtcamp8 vendors
 * (ext/stb_image_write.h; test-side checker, built only where the rtcamp8 tree is mounted).
 *
 *   stb_bmp_probe <width> <height> <seed> <rgba.bin> <out.bmp>
 * writes width*height pseudo-random RGBA8 pixels (every alpha value from 0 to 255 appears for images of
 * 256 pixels or more) to rgba.bin, and stbi_write_bmp(out.bmp, width, height, 4, pixels), the call
 * saveImage makes for .bmp (common/common_host.cpp:436-438). */
#define STB_IMAGE_WRITE_IMPLEMENTATION
#include "stb_image_write.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint32_t mix(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

int main(int argc, char** argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: stb_bmp_probe w h seed rgba.bin out.bmp\n");
        return 1;
    }
    int w = atoi(argv[1]), h = atoi(argv[2]);
    uint32_t seed = (uint32_t)strtoul(argv[3], NULL, 10);
    uint32_t* px = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)w * h);
    for (int i = 0; i < w * h; ++i) {
        uint32_t v = mix((uint32_t)i * 2654435761u ^ seed);
        uint32_t a = (uint32_t)i & 255u;   /* all alpha values */
        px[i] = (v & 0x00ffffffu) | (a << 24);
    }
    FILE* f = fopen(argv[4], "wb");
    if (!f) return 1;
    fwrite(px, 4, (size_t)w * h, f);
    fclose(f);
    if (!stbi_write_bmp(argv[5], w, h, 4, px)) return 1;
    free(px);
    return 0;
}

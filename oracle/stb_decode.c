/* Decodes image files with the reference's vendored stb_image 2.30
 * (ThirdParty/stb_image/stb_image.h) exactly as Texture::Load does
 * (Src/Core/texture.cu:207-214): stbi_load(path, &w, &h, &c, 4).
 *
 * TEST INFRASTRUCTURE ONLY: `make -C oracle ref` compiles this file against
 * the header where it lies in the reference tree, into oracle/_ref/; the
 * fixture script under tests/golden/ runs it to record expected texels.
 *
 * Usage: stb_decode IN OUT -- OUT receives int32 width, int32 height, then
 * width * height * 4 bytes (RGBA8, row 0 = top).  Exit status 2 and stbi's
 * failure reason on stderr when decoding fails. */
#define STB_IMAGE_IMPLEMENTATION
#include "stb_image.h"

#include <stdint.h>
#include <stdio.h>

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 1;
    }
    int w = 0, h = 0, c = 0;
    unsigned char* px = stbi_load(argv[1], &w, &h, &c, 4);
    if (!px) {
        fprintf(stderr, "%s\n", stbi_failure_reason());
        return 2;
    }
    FILE* f = fopen(argv[2], "wb");
    if (!f) return 1;
    int32_t hdr[2] = {w, h};
    int ok = fwrite(hdr, sizeof hdr, 1, f) == 1 && fwrite(px, (size_t)w * h * 4, 1, f) == 1;
    ok = fclose(f) == 0 && ok;
    stbi_image_free(px);
    return ok ? 0 : 1;
}

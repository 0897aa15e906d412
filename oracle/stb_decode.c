/* Decodes image files with the reference's vendored stb_image 2.30
 * (ThirdParty/stb_image/stb_image.h) exactly as the reference's loaders call it:
 *  - Texture::Load (Src/Core/texture.cu:207-214): stbi_load(path, &w, &h, &c, 4);
 *  - HDRTexture::Load (Src/Core/texture.cu:17-25): stbi_loadf(path, &w, &h, &c, 4).
 *
 * TEST INFRASTRUCTURE ONLY: `make -C oracle ref` compiles this file against
 * the header where it lies in the reference tree, into oracle/_ref/; the
 * fixture scripts under tests/golden/ run it to record expected texels.
 *
 * Usage: stb_decode [-f] IN OUT -- OUT receives int32 width, int32 height, then
 * width * height * 4 values, row 0 = top: RGBA8 bytes, or with -f float32
 * RGBA from stbi_loadf.  Exit status 2 and stbi's failure reason on stderr
 * when decoding fails. */
#define STB_IMAGE_IMPLEMENTATION
#include "stb_image.h"

#include <stdint.h>
#include <stdio.h>
#include <string.h>

int main(int argc, char** argv) {
    int as_float = argc == 4 && strcmp(argv[1], "-f") == 0;
    if (argc != 3 && !as_float) {
        fprintf(stderr, "usage: %s [-f] IN OUT\n", argv[0]);
        return 1;
    }
    const char* in = argv[argc - 2];
    const char* out = argv[argc - 1];
    int w = 0, h = 0, c = 0;
    void* px = as_float ? (void*)stbi_loadf(in, &w, &h, &c, 4) : (void*)stbi_load(in, &w, &h, &c, 4);
    if (!px) {
        fprintf(stderr, "%s\n", stbi_failure_reason());
        return 2;
    }
    FILE* f = fopen(out, "wb");
    if (!f) return 1;
    int32_t hdr[2] = {w, h};
    size_t bytes = (size_t)w * h * 4 * (as_float ? sizeof(float) : 1);
    int ok = fwrite(hdr, sizeof hdr, 1, f) == 1 && (bytes == 0 || fwrite(px, bytes, 1, f) == 1);
    ok = fclose(f) == 0 && ok;
    stbi_image_free(px);
    return ok ? 0 : 1;
}

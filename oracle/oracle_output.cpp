// Oracle (test infrastructure only): the reference's output drivers restated in plain C++.
//
//   image/hdr/hdr.go:24-48        convertRGBToRGBE
//   image/hdr/writer.go:62-107    Writer.WriteImage (header, scanlines bottom row first, no RLE)
//   builtin/driver/outputfloat.go:30-42   binary.Write(LittleEndian, FrameBuf())
//
// Go converts float32 to byte through a truncating int32 conversion (CVTTSS2SL on amd64, the machine
// the reference's assembly targets); _mm_cvtt_ss2si is that instruction, so NaN and out-of-range
// values give 0x80000000 and the low byte 0 as in the reference.
#include <xmmintrin.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

uint8_t go_byte(float x) { return (uint8_t)(uint32_t)_mm_cvtt_ss2si(_mm_set_ss(x)); }

void rgbe(float r, float g, float b, uint8_t* out) {
    float d = r;
    if (g > d) d = g;
    if (b > d) d = b;
    if (d < 0.000001f) { out[0] = out[1] = out[2] = out[3] = 0; return; }
    int e = 0;
    double nd = d;
    if (std::isfinite(d)) nd = std::frexp((double)d, &e);   // math.Frexp: (d, 0) for Inf and NaN
    const float n = (float)nd;
    const float df = n * 255.999f / d;
    out[0] = go_byte(r * df);
    out[1] = go_byte(g * df);
    out[2] = go_byte(b * df);
    out[3] = (uint8_t)(e + 128);
}

}  // namespace

extern "C" {

// RGBE bytes of a width x height float RGB film, bottom row first (writer.go:96-106).
void orc_film_to_rgbe(const float* film, int32_t w, int32_t h, uint8_t* out) {
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++) {
            const int k = h - j - 1;
            const float* p = film + ((size_t)i + (size_t)k * w) * 3;
            rgbe(p[0], p[1], p[2], out + ((size_t)j * w + i) * 4);
        }
}

// The whole .hdr file (writer.go:77-81 header, then the scanlines).  Returns 0 on success.
int orc_write_hdr(const float* film, int32_t w, int32_t h, const char* path) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    std::fprintf(f, "#?RADIANCE\n");
    std::fprintf(f, "# %s\n", "Created by Vermeer Light Tools (http://www.vermeerlt.com)");
    std::fprintf(f, "FORMAT=32-bit_rle_rgbe\n");
    std::fprintf(f, "\n");
    std::fprintf(f, "+Y %d +X %d\n", h, w);
    std::vector<uint8_t> line((size_t)w * 4);
    for (int j = 0; j < h; j++) {
        const int k = h - j - 1;
        for (int i = 0; i < w; i++) {
            const float* p = film + ((size_t)i + (size_t)k * w) * 3;
            rgbe(p[0], p[1], p[2], &line[(size_t)i * 4]);
        }
        std::fwrite(line.data(), 1, line.size(), f);
    }
    return std::fclose(f) == 0 ? 0 : -1;
}

// binary.Write(fp, binary.LittleEndian, film): the float32 values in order, little-endian.
int orc_write_float(const float* film, int32_t w, int32_t h, const char* path) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    for (size_t q = 0; q < (size_t)w * h * 3; q++) {
        uint32_t bits;
        std::memcpy(&bits, &film[q], 4);
        const uint8_t le[4] = {(uint8_t)bits, (uint8_t)(bits >> 8), (uint8_t)(bits >> 16), (uint8_t)(bits >> 24)};
        std::fwrite(le, 1, 4, f);
    }
    return std::fclose(f) == 0 ? 0 : -1;
}

}  // extern "C"

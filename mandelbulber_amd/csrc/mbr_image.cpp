// mbr_image.cpp — image compile and writers of the render's output.
//
// mbr_compile_image restates cImage::CalculatePixel / CompileImage
// (src/cimage.cpp:132-213): brightness, contrast, optional tanh HDR curve,
// clamp, 16-bit quantisation and the gamma table.  mbr_write_png writes the
// 16-bit image (and the alpha plane) as 8- or 16-bit PNG, as SavePNG /
// SavePNG16 / SavePNG16Alpha do (src/files.cpp), with zlib.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mbr.h"

namespace {

void put32(std::vector<unsigned char> &v, uint32_t x) {
  v.push_back((unsigned char)(x >> 24));
  v.push_back((unsigned char)(x >> 16));
  v.push_back((unsigned char)(x >> 8));
  v.push_back((unsigned char)x);
}

void chunk(FILE *f, const char *type, const std::vector<unsigned char> &data) {
  std::vector<unsigned char> head;
  put32(head, (uint32_t)data.size());
  head.insert(head.end(), type, type + 4);
  uLong crc = crc32(0L, Z_NULL, 0);
  crc = crc32(crc, head.data() + 4, 4);
  if (!data.empty()) crc = crc32(crc, data.data(), (uInt)data.size());
  std::vector<unsigned char> tail;
  put32(tail, (uint32_t)crc);
  std::fwrite(head.data(), 1, head.size(), f);
  if (!data.empty()) std::fwrite(data.data(), 1, data.size(), f);
  std::fwrite(tail.data(), 1, tail.size(), f);
}

}  // namespace

extern "C" {

int mbr_compile_image(const mbr_scene *s, int width, int height, const float *const rgb[3],
                      uint16_t *image16) {
  if (!s || !rgb || !rgb[0] || !rgb[1] || !rgb[2] || !image16 || width <= 0 || height <= 0)
    return MBR_ERR_INVALID_ARG;
  // CalculateGammaTable (src/cimage.cpp:176-182)
  std::vector<uint16_t> gamma(65536);
  for (int i = 0; i < 65536; i++)
    gamma[i] = (uint16_t)(int)(std::pow(i / 65536.0, 1.0 / s->image_gamma) * 65535);
  const size_t n = (size_t)width * height;
  for (size_t i = 0; i < n; i++) {
    float c[3];
    for (int k = 0; k < 3; k++) {
      // float channel times the double brightness, rounded back to float
      // (sImageAdjustments::brightness is a double, src/cimage.hpp:54)
      float v = (float)(rgb[k][i] * s->brightness);
      v = (float)((v - 0.5) * s->contrast + 0.5);
      if (v < 0.0f) v = 0.0f;
      if (s->hdr) v = std::tanh(v);
      if (v > 1.0f) v = 1.0f;
      c[k] = v;
    }
    for (int k = 0; k < 3; k++) image16[3 * i + k] = gamma[(uint16_t)(c[k] * 65535.0)];
  }
  return MBR_OK;
}

int mbr_write_png(const char *path, int width, int height, const uint16_t *image16, const uint16_t *alpha,
                  int bits) {
  if (!path || !image16 || width <= 0 || height <= 0 || (bits != 8 && bits != 16)) return MBR_ERR_INVALID_ARG;
  const int channels = alpha ? 4 : 3;
  const int bpp = channels * bits / 8;
  const size_t row = (size_t)width * bpp + 1;
  std::vector<unsigned char> raw(row * height);
  for (int y = 0; y < height; y++) {
    unsigned char *o = &raw[row * y];
    *o++ = 0;  // filter: none
    for (int x = 0; x < width; x++) {
      const size_t i = (size_t)y * width + x;
      uint16_t px[4] = {image16[3 * i], image16[3 * i + 1], image16[3 * i + 2], alpha ? alpha[i] : (uint16_t)0};
      for (int k = 0; k < channels; k++) {
        if (bits == 16) {
          *o++ = (unsigned char)(px[k] >> 8);
          *o++ = (unsigned char)(px[k] & 255);
        } else {
          *o++ = (unsigned char)(px[k] >> 8);
        }
      }
    }
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return MBR_ERR_IO;
  z.resize(zlen);
  FILE *f = std::fopen(path, "wb");
  if (!f) return MBR_ERR_IO;
  static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  std::fwrite(sig, 1, 8, f);
  std::vector<unsigned char> ihdr;
  put32(ihdr, (uint32_t)width);
  put32(ihdr, (uint32_t)height);
  ihdr.push_back((unsigned char)bits);
  ihdr.push_back(alpha ? 6 : 2);  // RGBA or RGB
  ihdr.push_back(0);
  ihdr.push_back(0);
  ihdr.push_back(0);
  chunk(f, "IHDR", ihdr);
  chunk(f, "IDAT", z);
  chunk(f, "IEND", {});
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? MBR_OK : MBR_ERR_IO;
}

}  // extern "C"

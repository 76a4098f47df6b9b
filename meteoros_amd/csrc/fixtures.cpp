// fixtures.cpp -- loaders for the noise volumes and maps Meteoros ships
// (src/CloudScapes/textures/CloudTextures/), so the stage can render the
// reference's own clouds.
//
//   mtc_load_tga_slices   ImageLoadingUtility.cpp:75-139: slice z of an
//                         N^3 volume is "<dir>/<base>(z+1).tga", decoded to
//                         RGBA8 and stored at z * N * N * 4, x fastest
//   mtc_load_image_rgba8  Texture2D::createTextureFromFile with 4 channels
//                         (Sky.cpp:44-57): curlNoise.png, weatherMap.png
//
// Decoding follows stbi_load(..., 4), which the reference calls:
//   TGA  types 2/10 (true colour, raw/RLE) and 3/11 (grey), 8/24/32 bits;
//        BGR(A) to RGBA, alpha 255 when absent; rows flipped unless the
//        descriptor's top-left origin bit is set
//   PNG  colour types 0, 2, 3, 4, 6 at 8 or 16 bits (16-bit samples keep
//        their high byte), the five scanline filters, zlib-compressed IDAT;
//        no Adam7 interlacing (MTC_ERR_UNSUPPORTED)
// Host code only; nothing here touches the GPU.
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "runtime_internal.h"

namespace {

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0) {
    std::fclose(f);
    return false;
  }
  out.resize((size_t)n);
  const bool ok = n == 0 || std::fread(out.data(), 1, (size_t)n, f) == (size_t)n;
  std::fclose(f);
  return ok;
}

struct Image8 {
  int w = 0, h = 0;
  std::vector<uint8_t> rgba;  // top row first
};

// Largest decoded image: 2^26 pixels (8192^2, 256 MiB of RGBA8), the largest
// map mtc_create accepts.  Sizes come from file headers, so they are checked
// against this bound and against what the file's bytes can hold before any
// buffer is allocated.
constexpr size_t kMaxPixels = size_t(1) << 26;

// ---------------------------------------------------------------------------
// TGA
// ---------------------------------------------------------------------------
int decode_tga(const std::vector<uint8_t>& d, Image8& img, std::string& err) {
  if (d.size() < 18) return err = "truncated header", MTC_ERR_INVALID_ARG;
  const int idlen = d[0], cmtype = d[1], type = d[2];
  const int w = d[12] | (d[13] << 8), h = d[14] | (d[15] << 8), bpp = d[16], desc = d[17];
  const bool rle = type == 10 || type == 11, grey = type == 3 || type == 11;
  if (cmtype != 0 || !(type == 2 || type == 3 || type == 10 || type == 11))
    return err = "TGA type " + std::to_string(type) + " not supported", MTC_ERR_UNSUPPORTED;
  if (w <= 0 || h <= 0) return err = "empty image", MTC_ERR_INVALID_ARG;
  if ((grey && bpp != 8) || (!grey && bpp != 24 && bpp != 32))
    return err = "TGA depth " + std::to_string(bpp) + " not supported", MTC_ERR_UNSUPPORTED;
  const int bytes = bpp / 8;
  size_t pos = 18 + (size_t)idlen;
  const size_t npx = (size_t)w * (size_t)h;
  if (npx > kMaxPixels) return err = "image larger than 2^26 pixels", MTC_ERR_INVALID_ARG;
  // raw data must be in the file; an RLE packet (1 + bytes) covers at most 128 pixels
  const size_t avail = d.size() > pos ? d.size() - pos : 0;
  if (!rle && npx * bytes > avail) return err = "truncated pixel data", MTC_ERR_INVALID_ARG;
  if (rle && npx > 128 * avail) return err = "truncated RLE data", MTC_ERR_INVALID_ARG;
  std::vector<uint8_t> px(npx * bytes);
  size_t o = 0;
  while (o < px.size()) {
    if (!rle) {
      if (pos + px.size() > d.size()) return err = "truncated pixel data", MTC_ERR_INVALID_ARG;
      std::memcpy(px.data(), d.data() + pos, px.size());
      o = px.size();
      break;
    }
    if (pos >= d.size()) return err = "truncated RLE data", MTC_ERR_INVALID_ARG;
    const int hdr = d[pos++];
    const size_t count = (size_t)(hdr & 0x7f) + 1;
    if (o + count * bytes > px.size()) return err = "RLE run past the image", MTC_ERR_INVALID_ARG;
    if (hdr & 0x80) {
      if (pos + bytes > d.size()) return err = "truncated RLE data", MTC_ERR_INVALID_ARG;
      for (size_t k = 0; k < count; ++k) std::memcpy(px.data() + o + k * bytes, d.data() + pos, bytes);
      pos += bytes;
    } else {
      if (pos + count * bytes > d.size()) return err = "truncated RLE data", MTC_ERR_INVALID_ARG;
      std::memcpy(px.data() + o, d.data() + pos, count * bytes);
      pos += count * bytes;
    }
    o += count * bytes;
  }
  img.w = w;
  img.h = h;
  img.rgba.assign((size_t)w * h * 4, 0);
  const bool top_left = (desc >> 5) & 1;
  for (int y = 0; y < h; ++y) {
    const int sy = top_left ? y : h - 1 - y;
    for (int x = 0; x < w; ++x) {
      const uint8_t* s = px.data() + ((size_t)sy * w + x) * bytes;
      uint8_t* t = img.rgba.data() + ((size_t)y * w + x) * 4;
      if (grey) {
        t[0] = t[1] = t[2] = s[0];
        t[3] = 255;
      } else {
        t[0] = s[2];
        t[1] = s[1];
        t[2] = s[0];
        t[3] = bytes == 4 ? s[3] : 255;
      }
    }
  }
  return MTC_OK;
}

// ---------------------------------------------------------------------------
// PNG
// ---------------------------------------------------------------------------
uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

int paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

int decode_png(const std::vector<uint8_t>& d, Image8& img, std::string& err) {
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  if (d.size() < 8 || std::memcmp(d.data(), sig, 8) != 0) return err = "not a PNG file", MTC_ERR_INVALID_ARG;
  size_t pos = 8;
  int w = 0, h = 0, depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, palette, trns;
  while (pos + 8 <= d.size()) {
    const uint32_t len = be32(d.data() + pos);
    const std::string type((const char*)d.data() + pos + 4, 4);
    if (pos + 12 + (size_t)len > d.size()) return err = "truncated chunk", MTC_ERR_INVALID_ARG;
    const uint8_t* c = d.data() + pos + 8;
    if (type == "IHDR") {
      if (len < 13) return err = "bad IHDR", MTC_ERR_INVALID_ARG;
      w = (int)be32(c);
      h = (int)be32(c + 4);
      depth = c[8];
      ctype = c[9];
      interlace = c[12];
    } else if (type == "PLTE") {
      palette.assign(c, c + len);
    } else if (type == "tRNS") {
      trns.assign(c, c + len);
    } else if (type == "IDAT") {
      idat.insert(idat.end(), c, c + len);
    } else if (type == "IEND") {
      break;
    }
    pos += 12 + (size_t)len;
  }
  if (w <= 0 || h <= 0 || w > (1 << 16) || h > (1 << 16)) return err = "bad PNG size", MTC_ERR_INVALID_ARG;
  if ((size_t)w * (size_t)h > kMaxPixels) return err = "image larger than 2^26 pixels", MTC_ERR_INVALID_ARG;
  if (interlace) return err = "interlaced PNG not supported", MTC_ERR_UNSUPPORTED;
  int channels;
  switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: return err = "bad PNG colour type", MTC_ERR_INVALID_ARG;
  }
  if (!(depth == 8 || (depth == 16 && ctype != 3)))
    return err = "PNG bit depth " + std::to_string(depth) + " not supported", MTC_ERR_UNSUPPORTED;
  if (ctype == 3 && palette.empty()) return err = "palette image without PLTE", MTC_ERR_INVALID_ARG;
  const int bpp = channels * depth / 8;  // bytes per pixel
  const size_t stride = (size_t)w * bpp;
  // deflate expands at most 1032:1, so a short IDAT cannot hold a large image
  if ((stride + 1) * (size_t)h > 1032 * idat.size() + 64) return err = "truncated IDAT stream", MTC_ERR_INVALID_ARG;
  std::vector<uint8_t> raw((stride + 1) * (size_t)h);
  uLongf rawlen = (uLongf)raw.size();
  if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK || rawlen != raw.size())
    return err = "corrupt IDAT stream", MTC_ERR_INVALID_ARG;
  std::vector<uint8_t> px(stride * h);
  for (int y = 0; y < h; ++y) {
    const int filter = raw[(stride + 1) * y];
    const uint8_t* in = raw.data() + (stride + 1) * y + 1;
    uint8_t* out = px.data() + stride * y;
    const uint8_t* up = y ? px.data() + stride * (y - 1) : nullptr;
    for (size_t i = 0; i < stride; ++i) {
      const int a = i >= (size_t)bpp ? out[i - bpp] : 0;
      const int b = up ? up[i] : 0;
      const int cc = (up && i >= (size_t)bpp) ? up[i - bpp] : 0;
      int v = in[i];
      switch (filter) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, cc); break;
        default: return err = "bad scanline filter", MTC_ERR_INVALID_ARG;
      }
      out[i] = (uint8_t)v;
    }
  }
  img.w = w;
  img.h = h;
  img.rgba.assign((size_t)w * h * 4, 255);
  const int sb = depth / 8;  // bytes per sample; 16-bit samples keep the high byte
  for (size_t p = 0; p < (size_t)w * h; ++p) {
    const uint8_t* s = px.data() + p * bpp;
    uint8_t* t = img.rgba.data() + p * 4;
    switch (ctype) {
      case 0: t[0] = t[1] = t[2] = s[0]; break;
      case 4: t[0] = t[1] = t[2] = s[0]; t[3] = s[sb]; break;
      case 2: t[0] = s[0]; t[1] = s[sb]; t[2] = s[2 * sb]; break;
      case 6: t[0] = s[0]; t[1] = s[sb]; t[2] = s[2 * sb]; t[3] = s[3 * sb]; break;
      case 3: {
        const size_t k = s[0];
        if (3 * k + 2 >= palette.size()) return err = "palette index out of range", MTC_ERR_INVALID_ARG;
        t[0] = palette[3 * k];
        t[1] = palette[3 * k + 1];
        t[2] = palette[3 * k + 2];
        t[3] = k < trns.size() ? trns[k] : 255;
        break;
      }
    }
  }
  return MTC_OK;
}

int decode_any(const std::string& path, Image8& img, std::string& err) {
  std::vector<uint8_t> d;
  if (!read_file(path, d)) return err = "cannot read " + path, MTC_ERR_INVALID_ARG;
  if (d.size() >= 8 && d[0] == 0x89 && d[1] == 'P') return decode_png(d, img, err);
  return decode_tga(d, img, err);
}

}  // namespace

extern "C" {

int mtc_load_image_rgba8(const char* path, int32_t* width, int32_t* height, uint8_t* out, size_t n_bytes) try {
  if (!path) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_load_image_rgba8: null path");
  Image8 img;
  std::string err;
  int rc = decode_any(path, img, err);
  if (rc) return mtc_fail(rc, "mtc_load_image_rgba8: " + std::string(path) + ": " + err);
  if (width) *width = img.w;
  if (height) *height = img.h;
  if (out) {
    if (n_bytes < img.rgba.size()) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_load_image_rgba8: buffer too small");
    std::memcpy(out, img.rgba.data(), img.rgba.size());
  }
  return MTC_OK;
} catch (const std::bad_alloc&) {
  return mtc_fail(MTC_ERR_OUT_OF_MEMORY, "mtc_load_image_rgba8: out of host memory");
} catch (...) {
  return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_load_image_rgba8: decoding failed");
}

int mtc_load_tga_slices(const char* dir, const char* base, int32_t n, uint8_t* out, size_t n_bytes) try {
  if (!dir || !base || !out) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_load_tga_slices: null pointer");
  if (n < 1 || n > 1024) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_load_tga_slices: n must be in 1..1024");
  const size_t slice = (size_t)n * n * 4;
  if (n_bytes < slice * n) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_load_tga_slices: buffer smaller than n^3 * 4 bytes");
  for (int z = 0; z < n; ++z) {
    const std::string path = std::string(dir) + "/" + base + "(" + std::to_string(z + 1) + ").tga";
    Image8 img;
    std::string err;
    int rc = decode_any(path, img, err);
    if (rc) return mtc_fail(rc, "mtc_load_tga_slices: " + path + ": " + err);
    if (img.w != n || img.h != n) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_load_tga_slices: " + path + " is not n x n");
    std::memcpy(out + slice * z, img.rgba.data(), slice);
  }
  return MTC_OK;
} catch (const std::bad_alloc&) {
  return mtc_fail(MTC_ERR_OUT_OF_MEMORY, "mtc_load_tga_slices: out of host memory");
} catch (...) {
  return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_load_tga_slices: decoding failed");
}

}  // extern "C"

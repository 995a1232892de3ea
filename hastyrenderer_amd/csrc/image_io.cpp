// Image readers and the EXR writer (image_io.h).
#include "image_io.h"

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace hasty {

namespace {

std::vector<uint8_t> readAll(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("could not open file " + path + " for reading");
  return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

[[noreturn]] void bad(const std::string& path, const std::string& what) {
  throw std::runtime_error("could not read " + path + ": " + what);
}

// Image sizes the readers accept: at most 2^20 pixels a side and 2^30 in
// all, so that no size product below can overflow.
void checkDims(const std::string& path, uint64_t W, uint64_t H) {
  if (W == 0 || H == 0 || W > (1u << 20) || H > (1u << 20) || W * H > (1ull << 30))
    bad(path, "image size " + std::to_string(W) + "x" + std::to_string(H) + " out of range");
}

// ------------------------------------------------------------- inflate ----
// Deflate decoder written from RFC 1951.  Codes are canonical Huffman codes:
// for each code length the number of codes and the symbols in code order are
// kept, and a symbol is decoded by reading bits MSB-first of the code while
// walking the lengths (codes of one length are consecutive integers).
struct BitReader {
  const uint8_t* p;
  size_t n, pos = 0;  // byte position
  uint32_t bitbuf = 0;
  int bitcnt = 0;
  BitReader(const uint8_t* d, size_t size) : p(d), n(size) {}
  uint32_t bits(int need) {  // LSB-first, need <= 24
    while (bitcnt < need) {
      if (pos >= n) throw std::runtime_error("deflate stream ends early");
      bitbuf |= (uint32_t)p[pos++] << bitcnt;
      bitcnt += 8;
    }
    uint32_t v = bitbuf & ((1u << need) - 1u);
    bitbuf >>= need;
    bitcnt -= need;
    return v;
  }
  void alignByte() {
    bitbuf = 0;
    bitcnt = 0;
  }
};

struct Huffman {
  uint16_t count[16] = {0};  // codes of each length
  std::vector<uint16_t> symbol;
  // lengths[i] = code length of symbol i (0 = unused); false if the set of
  // lengths over-subscribes the code space
  bool build(const uint8_t* lengths, int n) {
    memset(count, 0, sizeof(count));
    for (int i = 0; i < n; ++i) count[lengths[i]]++;
    count[0] = 0;
    int left = 1;
    for (int len = 1; len < 16; ++len) {
      left <<= 1;
      left -= count[len];
      if (left < 0) return false;
    }
    uint16_t offs[16];
    offs[1] = 0;
    for (int len = 1; len < 15; ++len) offs[len + 1] = offs[len] + count[len];
    symbol.assign(n, 0);
    for (int i = 0; i < n; ++i)
      if (lengths[i]) symbol[offs[lengths[i]]++] = (uint16_t)i;
    return true;
  }
  int decode(BitReader& br) const {
    int code = 0, first = 0, index = 0;
    for (int len = 1; len < 16; ++len) {
      code |= (int)br.bits(1);
      const int c = count[len];
      if (code - c < first) return symbol[index + (code - first)];
      index += c;
      first += c;
      first <<= 1;
      code <<= 1;
    }
    throw std::runtime_error("invalid Huffman code in deflate stream");
  }
};

const uint16_t kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                               35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

void inflateBlock(BitReader& br, std::vector<uint8_t>& out, const Huffman& lit, const Huffman& dist,
                  size_t maxOut) {
  for (;;) {
    if (out.size() > maxOut) throw std::runtime_error("deflate stream longer than expected");
    int sym = lit.decode(br);
    if (sym < 256) {
      out.push_back((uint8_t)sym);
    } else if (sym == 256) {
      return;
    } else {
      sym -= 257;
      if (sym >= 29) throw std::runtime_error("invalid length code in deflate stream");
      const size_t len = kLenBase[sym] + br.bits(kLenExtra[sym]);
      const int ds = dist.decode(br);
      if (ds >= 30) throw std::runtime_error("invalid distance code in deflate stream");
      const size_t d = kDistBase[ds] + br.bits(kDistExtra[ds]);
      if (d > out.size()) throw std::runtime_error("deflate distance beyond the output");
      const size_t from = out.size() - d;
      for (size_t i = 0; i < len; ++i) out.push_back(out[from + i]);  // may overlap: byte by byte
    }
  }
}

}  // namespace

std::vector<uint8_t> zlibInflate(const uint8_t* data, size_t size, size_t maxOut) {
  if (size < 2) throw std::runtime_error("zlib stream too short");
  const uint8_t cmf = data[0], flg = data[1];
  if ((cmf & 0x0F) != 8 || ((cmf << 8) | flg) % 31 != 0) throw std::runtime_error("not a zlib deflate stream");
  if (flg & 0x20) throw std::runtime_error("zlib preset dictionaries are not supported");
  BitReader br(data + 2, size - 2);
  std::vector<uint8_t> out;
  Huffman fixedLit, fixedDist;
  {
    uint8_t l[288];
    for (int i = 0; i < 144; ++i) l[i] = 8;
    for (int i = 144; i < 256; ++i) l[i] = 9;
    for (int i = 256; i < 280; ++i) l[i] = 7;
    for (int i = 280; i < 288; ++i) l[i] = 8;
    fixedLit.build(l, 288);
    uint8_t d[30];
    for (int i = 0; i < 30; ++i) d[i] = 5;
    fixedDist.build(d, 30);
  }
  bool last = false;
  while (!last) {
    last = br.bits(1) != 0;
    const uint32_t type = br.bits(2);
    if (type == 0) {  // stored
      br.alignByte();
      if (br.pos + 4 > br.n) throw std::runtime_error("deflate stream ends early");
      const uint32_t len = br.p[br.pos] | (br.p[br.pos + 1] << 8);
      const uint32_t nlen = br.p[br.pos + 2] | (br.p[br.pos + 3] << 8);
      br.pos += 4;
      if ((len ^ 0xFFFFu) != nlen) throw std::runtime_error("corrupt stored deflate block");
      if (br.pos + len > br.n) throw std::runtime_error("deflate stream ends early");
      if (out.size() + len > maxOut) throw std::runtime_error("deflate stream longer than expected");
      out.insert(out.end(), br.p + br.pos, br.p + br.pos + len);
      br.pos += len;
    } else if (type == 1) {
      inflateBlock(br, out, fixedLit, fixedDist, maxOut);
    } else if (type == 2) {
      const int nlen = (int)br.bits(5) + 257, ndist = (int)br.bits(5) + 1, ncode = (int)br.bits(4) + 4;
      if (nlen > 286 || ndist > 30) throw std::runtime_error("bad deflate code counts");
      static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
      uint8_t clen[19] = {0};
      for (int i = 0; i < ncode; ++i) clen[order[i]] = (uint8_t)br.bits(3);
      Huffman lencode;
      if (!lencode.build(clen, 19)) throw std::runtime_error("bad deflate code-length code");
      uint8_t lengths[320] = {0};
      int i = 0;
      while (i < nlen + ndist) {
        int sym = lencode.decode(br);
        if (sym < 16) {
          lengths[i++] = (uint8_t)sym;
        } else {
          uint8_t val = 0;
          int rep;
          if (sym == 16) {
            if (i == 0) throw std::runtime_error("deflate repeat with no previous length");
            val = lengths[i - 1];
            rep = 3 + (int)br.bits(2);
          } else if (sym == 17) {
            rep = 3 + (int)br.bits(3);
          } else {
            rep = 11 + (int)br.bits(7);
          }
          if (i + rep > nlen + ndist) throw std::runtime_error("deflate code lengths overflow");
          while (rep--) lengths[i++] = val;
        }
      }
      Huffman lit, dist;
      if (!lit.build(lengths, nlen) || !dist.build(lengths + nlen, ndist))
        throw std::runtime_error("bad deflate literal or distance code");
      inflateBlock(br, out, lit, dist, maxOut);
    } else {
      throw std::runtime_error("invalid deflate block type");
    }
  }
  return out;
}

namespace {

// ----------------------------------------------------------------- PNG ----
uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

FloatImage readPNG(const std::string& path, const std::vector<uint8_t>& f) {
  size_t pos = 8;
  uint32_t W = 0, H = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, palette;
  bool ended = false;
  while (pos + 8 <= f.size() && !ended) {
    const uint32_t len = be32(&f[pos]);
    const char* type = (const char*)&f[pos + 4];
    if (pos + 12 + (size_t)len > f.size()) bad(path, "truncated PNG chunk");
    const uint8_t* d = &f[pos + 8];
    if (!memcmp(type, "IHDR", 4)) {
      if (len < 13) bad(path, "short IHDR");
      W = be32(d);
      H = be32(d + 4);
      depth = d[8];
      ctype = d[9];
      interlace = d[12];
    } else if (!memcmp(type, "PLTE", 4)) {
      palette.assign(d, d + len);
    } else if (!memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), d, d + len);
    } else if (!memcmp(type, "IEND", 4)) {
      ended = true;
    }
    pos += 12 + (size_t)len;
  }
  if (W == 0 || H == 0) bad(path, "PNG without image header");
  checkDims(path, W, H);
  if (interlace) bad(path, "interlaced PNG images are not supported");
  int samples;
  switch (ctype) {
    case 0: samples = 1; break;
    case 2: samples = 3; break;
    case 3: samples = 1; break;
    case 4: samples = 2; break;
    case 6: samples = 4; break;
    default: bad(path, "unknown PNG colour type");
  }
  const bool okDepth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                       (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                       ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
  if (!okDepth) bad(path, "unsupported PNG bit depth");
  if (ctype == 3 && palette.empty()) bad(path, "palette PNG without palette");
  const size_t bitsPerPixel = (size_t)samples * depth;
  const size_t stride = (W * bitsPerPixel + 7) / 8;
  const size_t bpp = bitsPerPixel >= 8 ? bitsPerPixel / 8 : 1;
  std::vector<uint8_t> raw = zlibInflate(idat.data(), idat.size(), (stride + 1) * (size_t)H);
  if (raw.size() < (stride + 1) * (size_t)H) bad(path, "PNG image data too short");
  // undo the row filters in place
  std::vector<uint8_t> img(stride * H);
  for (uint32_t y = 0; y < H; ++y) {
    const uint8_t ft = raw[y * (stride + 1)];
    const uint8_t* src = &raw[y * (stride + 1) + 1];
    uint8_t* row = &img[y * stride];
    const uint8_t* up = y ? &img[(y - 1) * stride] : nullptr;
    for (size_t x = 0; x < stride; ++x) {
      const int a = x >= bpp ? row[x - bpp] : 0;
      const int b = up ? up[x] : 0;
      const int c = (up && x >= bpp) ? up[x - bpp] : 0;
      int v;
      switch (ft) {
        case 0: v = 0; break;
        case 1: v = a; break;
        case 2: v = b; break;
        case 3: v = (a + b) / 2; break;
        case 4: {
          const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
          v = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
          break;
        }
        default: bad(path, "unknown PNG row filter");
      }
      row[x] = (uint8_t)(src[x] + v);
    }
  }
  FloatImage out;
  out.width = W;
  out.height = H;
  out.channels = ctype == 3 ? 3u : (uint32_t)samples;
  out.px.resize((size_t)W * H * out.channels);
  const float maxv = depth == 16 ? 65535.0f : (float)((1 << depth) - 1);
  for (uint32_t y = 0; y < H; ++y) {
    const uint8_t* row = &img[y * stride];
    for (uint32_t x = 0; x < W; ++x) {
      float* o = &out.px[((size_t)y * W + x) * out.channels];
      if (depth == 16) {
        for (int s = 0; s < samples; ++s) {
          const uint8_t* q = row + ((size_t)x * samples + s) * 2;
          o[s] = (float)((q[0] << 8) | q[1]) / maxv;
        }
      } else if (depth == 8) {
        if (ctype == 3) {
          const size_t k = row[x];
          if (3 * k + 2 >= palette.size()) bad(path, "PNG palette index out of range");
          for (int s = 0; s < 3; ++s) o[s] = (float)palette[3 * k + s] / 255.0f;
        } else {
          for (int s = 0; s < samples; ++s) o[s] = (float)row[(size_t)x * samples + s] / maxv;
        }
      } else {  // 1, 2, 4 bits: one sample per pixel, MSB first
        const size_t bit = (size_t)x * depth;
        const int v = (row[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
        if (ctype == 3) {
          if (3 * (size_t)v + 2 >= palette.size()) bad(path, "PNG palette index out of range");
          for (int s = 0; s < 3; ++s) o[s] = (float)palette[3 * v + s] / 255.0f;
        } else {
          o[0] = (float)v / maxv;
        }
      }
    }
  }
  return out;
}

// ------------------------------------------------------------- PPM / PGM ----
FloatImage readPNM(const std::string& path, const std::vector<uint8_t>& f) {
  const char kind = (char)f[1];
  const bool ascii = kind == '2' || kind == '3';
  const uint32_t ch = (kind == '3' || kind == '6') ? 3u : 1u;
  size_t pos = 2;
  auto token = [&]() -> long {
    for (;;) {
      while (pos < f.size() && isspace(f[pos])) ++pos;
      if (pos < f.size() && f[pos] == '#') {
        while (pos < f.size() && f[pos] != '\n') ++pos;
        continue;
      }
      break;
    }
    if (pos >= f.size() || !isdigit(f[pos])) bad(path, "malformed PNM header");
    long v = 0;
    while (pos < f.size() && isdigit(f[pos])) {
      v = v * 10 + (f[pos++] - '0');
      if (v > (1l << 30)) bad(path, "PNM number out of range");
    }
    return v;
  };
  const long W = token(), H = token(), maxval = token();
  if (W <= 0 || H <= 0 || maxval <= 0 || maxval > 65535) bad(path, "malformed PNM header");
  checkDims(path, (uint64_t)W, (uint64_t)H);
  FloatImage out;
  out.width = (uint32_t)W;
  out.height = (uint32_t)H;
  out.channels = ch;
  const size_t n = (size_t)W * H * ch;
  out.px.resize(n);
  if (ascii) {
    for (size_t i = 0; i < n; ++i) out.px[i] = (float)token() / (float)maxval;
  } else {
    ++pos;  // the single whitespace after maxval
    const size_t bytes = maxval > 255 ? 2 : 1;
    if (pos + n * bytes > f.size()) bad(path, "truncated PNM data");
    for (size_t i = 0; i < n; ++i) {
      const uint32_t v = bytes == 2 ? ((uint32_t)f[pos + 2 * i] << 8) | f[pos + 2 * i + 1] : f[pos + i];
      out.px[i] = (float)v / (float)maxval;
    }
  }
  return out;
}

// ----------------------------------------------------------------- PFM ----
FloatImage readPFMAny(const std::string& path, const std::vector<uint8_t>& f) {
  const uint32_t ch = f[1] == 'F' ? 3u : 1u;
  size_t pos = 2;
  auto word = [&]() {
    while (pos < f.size() && isspace(f[pos])) ++pos;
    size_t s = pos;
    while (pos < f.size() && !isspace(f[pos])) ++pos;
    return std::string((const char*)&f[s], pos - s);
  };
  const long W = atol(word().c_str()), H = atol(word().c_str());
  const double scale = atof(word().c_str());
  ++pos;
  if (W <= 0 || H <= 0 || scale == 0.0) bad(path, "malformed PFM header");
  checkDims(path, (uint64_t)W, (uint64_t)H);
  const size_t n = (size_t)W * H * ch;
  if (pos + n * 4 > f.size()) bad(path, "truncated PFM data");
  FloatImage out;
  out.width = (uint32_t)W;
  out.height = (uint32_t)H;
  out.channels = ch;
  out.px.resize(n);
  const bool little = scale < 0.0;
  for (long y = 0; y < H; ++y) {  // the file stores the bottom row first
    const uint8_t* src = &f[pos + (size_t)(H - 1 - y) * W * ch * 4];
    for (size_t i = 0; i < (size_t)W * ch; ++i) {
      uint8_t b[4];
      for (int k = 0; k < 4; ++k) b[k] = little ? src[4 * i + k] : src[4 * i + 3 - k];
      memcpy(&out.px[(size_t)y * W * ch + i], b, 4);
    }
  }
  return out;
}

// ----------------------------------------------------------------- EXR ----
// IEEE half -> float, exact for every half value (normals, subnormals,
// infinities, NaN payloads).
float halfToFloat(uint16_t h) {
  const uint32_t sign = (uint32_t)(h >> 15) << 31;
  const uint32_t exp = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
  uint32_t bits;
  if (exp == 0x1Fu) {
    bits = sign | 0x7F800000u | (man << 13);
  } else if (exp != 0) {
    bits = sign | ((exp + 112u) << 23) | (man << 13);
  } else if (man == 0) {
    bits = sign;
  } else {  // subnormal: man * 2^-24
    float v = (float)man * (1.0f / 16777216.0f);
    memcpy(&bits, &v, 4);
    bits |= sign;
  }
  float out;
  memcpy(&out, &bits, 4);
  return out;
}

uint32_t le32(const uint8_t* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }

// Undoes the byte predictor and the even/odd byte split that the RLE and ZIP
// compressors apply before coding (OpenEXR file layout, "ZIP compression").
void exrUnpredict(std::vector<uint8_t>& t) {
  for (size_t i = 1; i < t.size(); ++i) t[i] = (uint8_t)(t[i - 1] + t[i] - 128);
  std::vector<uint8_t> out(t.size());
  const size_t half = (t.size() + 1) / 2;
  for (size_t i = 0; i < t.size(); ++i) out[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
  t.swap(out);
}

std::vector<uint8_t> exrRleDecode(const std::string& path, const uint8_t* p, size_t n, size_t want) {
  std::vector<uint8_t> out;
  size_t i = 0;
  while (i < n) {
    const int c = (int8_t)p[i++];
    if (c < 0) {
      const size_t len = (size_t)(-c);
      if (i + len > n) bad(path, "corrupt RLE data");
      out.insert(out.end(), p + i, p + i + len);
      i += len;
    } else {
      if (i >= n) bad(path, "corrupt RLE data");
      out.insert(out.end(), (size_t)c + 1, p[i++]);
    }
    if (out.size() > want) bad(path, "RLE data longer than the chunk");
  }
  return out;
}

FloatImage readEXR(const std::string& path, const std::vector<uint8_t>& f) {
  if (f.size() < 8) bad(path, "truncated EXR header");
  const uint32_t version = le32(&f[4]);
  if ((version & 0xFFu) != 2) bad(path, "unsupported EXR version");
  if (version & 0x1A00u) bad(path, "tiled, deep and multi-part EXR files are not supported");
  struct Channel {
    std::string name;
    int type;
  };
  std::vector<Channel> chans;
  int compression = -1;
  int32_t box[4] = {0, 0, -1, -1};
  size_t pos = 8;
  auto cstr = [&]() {
    size_t s = pos;
    while (pos < f.size() && f[pos]) ++pos;
    if (pos >= f.size()) bad(path, "truncated EXR header");
    return std::string((const char*)&f[s], pos++ - s);
  };
  for (;;) {
    if (pos >= f.size()) bad(path, "truncated EXR header");
    if (f[pos] == 0) {
      ++pos;
      break;
    }
    const std::string name = cstr(), type = cstr();
    if (pos + 4 > f.size()) bad(path, "truncated EXR header");
    const uint32_t size = le32(&f[pos]);
    pos += 4;
    if (pos + size > f.size()) bad(path, "truncated EXR header");
    const uint8_t* v = &f[pos];
    if (name == "channels" && type == "chlist") {
      size_t q = 0;
      while (q < size && v[q]) {
        size_t s = q;
        while (q < size && v[q]) ++q;
        if (q + 17 > size) bad(path, "bad channel list");
        Channel c;
        c.name.assign((const char*)v + s, q - s);
        ++q;
        c.type = (int)le32(v + q);
        const int32_t xs = (int32_t)le32(v + q + 8), ys = (int32_t)le32(v + q + 12);
        if (xs != 1 || ys != 1) bad(path, "subsampled EXR channels are not supported");
        if (c.type < 0 || c.type > 2) bad(path, "unknown EXR pixel type");
        q += 16;
        chans.push_back(c);
      }
    } else if (name == "compression" && size >= 1) {
      compression = v[0];
    } else if (name == "dataWindow" && size >= 16) {
      for (int k = 0; k < 4; ++k) box[k] = (int32_t)le32(v + 4 * k);
    }
    pos += size;
  }
  if (chans.empty()) bad(path, "EXR file without channels");
  const int64_t W = (int64_t)box[2] - box[0] + 1, H = (int64_t)box[3] - box[1] + 1;
  if (W <= 0 || H <= 0) bad(path, "bad EXR data window");
  checkDims(path, (uint64_t)W, (uint64_t)H);
  int linesPerChunk;
  switch (compression) {
    case 0: case 1: case 2: linesPerChunk = 1; break;  // NONE, RLE, ZIPS
    case 3: linesPerChunk = 16; break;                 // ZIP
    default: bad(path, "EXR compression method " + std::to_string(compression) +
                           " is not supported (NONE, RLE, ZIPS and ZIP are read)");
  }
  // Output channels: R, G, B, A when present (in that order), else the single
  // channel of a one-channel file (OIIO's order for RGBA files).
  std::vector<int> pick;
  for (const char* want : {"R", "G", "B", "A"})
    for (size_t c = 0; c < chans.size(); ++c)
      if (chans[c].name == want) pick.push_back((int)c);
  if (pick.empty() && chans.size() == 1) pick.push_back(0);
  if (pick.empty()) bad(path, "EXR file without R, G, B or A channels");
  size_t lineBytes = 0;
  std::vector<size_t> chanOffset(chans.size());
  for (size_t c = 0; c < chans.size(); ++c) {
    chanOffset[c] = lineBytes;
    lineBytes += (size_t)W * (chans[c].type == 1 ? 2 : 4);
  }
  const size_t nChunks = (size_t)((H + linesPerChunk - 1) / linesPerChunk);
  if (pos + 8 * nChunks > f.size()) bad(path, "truncated EXR offset table");
  FloatImage out;
  out.width = (uint32_t)W;
  out.height = (uint32_t)H;
  out.channels = (uint32_t)pick.size();
  out.px.assign((size_t)W * H * out.channels, 0.0f);
  std::vector<uint8_t> done((size_t)H, 0);
  for (size_t k = 0; k < nChunks; ++k) {
    uint64_t off;
    memcpy(&off, &f[pos + 8 * k], 8);
    if (off > f.size() - 8) bad(path, "EXR chunk offset out of range");  // f.size() >= 8 here
    const int32_t y0 = (int32_t)le32(&f[off]) - box[1];
    const uint32_t dsize = le32(&f[off + 4]);
    if (dsize > f.size() - off - 8) bad(path, "truncated EXR chunk");
    if (y0 < 0 || y0 >= H || y0 % linesPerChunk) bad(path, "bad EXR chunk line");
    const int lines = (int)std::min<int64_t>(linesPerChunk, H - y0);
    const size_t want = lineBytes * lines;
    const uint8_t* d = &f[off + 8];
    std::vector<uint8_t> raw;
    if (compression == 0 || dsize == want) {  // stored as is when coding would not shrink it
      if (dsize != want) bad(path, "EXR chunk size mismatch");
      raw.assign(d, d + dsize);
    } else {
      raw = compression == 1 ? exrRleDecode(path, d, dsize, want) : zlibInflate(d, dsize, want);
      if (raw.size() != want) bad(path, "EXR chunk decodes to the wrong size");
      exrUnpredict(raw);
    }
    for (int l = 0; l < lines; ++l) {
      const uint8_t* line = &raw[(size_t)l * lineBytes];
      done[y0 + l] = 1;
      for (size_t o = 0; o < pick.size(); ++o) {
        const Channel& c = chans[pick[o]];
        const uint8_t* src = line + chanOffset[pick[o]];
        float* dst = &out.px[((size_t)(y0 + l) * W) * out.channels + o];
        for (int64_t x = 0; x < W; ++x, dst += out.channels) {
          if (c.type == 1) {
            *dst = halfToFloat((uint16_t)(src[2 * x] | (src[2 * x + 1] << 8)));
          } else if (c.type == 2) {
            uint32_t b = le32(src + 4 * x);
            memcpy(dst, &b, 4);
          } else {
            *dst = (float)le32(src + 4 * x);
          }
        }
      }
    }
  }
  for (uint8_t v : done)
    if (!v) bad(path, "EXR file misses scanlines");
  return out;
}

}  // namespace

FloatImage readImage(const std::string& path) {
  std::vector<uint8_t> f = readAll(path);
  static const uint8_t pngSig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  if (f.size() >= 8 && !memcmp(f.data(), pngSig, 8)) return readPNG(path, f);
  if (f.size() >= 4 && le32(f.data()) == 20000630u) return readEXR(path, f);
  if (f.size() >= 2 && f[0] == 'P' && (f[1] == 'F' || f[1] == 'f')) return readPFMAny(path, f);
  if (f.size() >= 2 && f[0] == 'P' && (f[1] == '2' || f[1] == '3' || f[1] == '5' || f[1] == '6'))
    return readPNM(path, f);
  bad(path, "unsupported image format (PNG, PFM, PPM, PGM and EXR are read)");
}

// ----------------------------------------------------------------- EXR ----
void writeEXR(const std::string& path, uint32_t W, uint32_t H, const float* rgb) {
  std::vector<uint8_t> h;
  auto put = [&](const void* p, size_t n) {
    const uint8_t* b = (const uint8_t*)p;
    h.insert(h.end(), b, b + n);
  };
  auto putU32 = [&](uint32_t v) { put(&v, 4); };  // little-endian host
  auto putStr = [&](const char* s) { put(s, strlen(s) + 1); };
  auto attr = [&](const char* name, const char* type, const void* data, uint32_t size) {
    putStr(name);
    putStr(type);
    putU32(size);
    put(data, size);
  };
  putU32(20000630u);  // magic
  putU32(2u);         // version 2, single-part scanline
  {
    std::vector<uint8_t> cl;
    for (const char* c : {"B", "G", "R"}) {
      cl.insert(cl.end(), c, c + 2);
      const int32_t pixelType = 2;  // FLOAT
      const uint8_t linear[4] = {0, 0, 0, 0};
      const int32_t samp[2] = {1, 1};
      cl.insert(cl.end(), (const uint8_t*)&pixelType, (const uint8_t*)&pixelType + 4);
      cl.insert(cl.end(), linear, linear + 4);
      cl.insert(cl.end(), (const uint8_t*)samp, (const uint8_t*)samp + 8);
    }
    cl.push_back(0);
    attr("channels", "chlist", cl.data(), (uint32_t)cl.size());
  }
  const uint8_t none = 0;
  attr("compression", "compression", &none, 1);
  const int32_t box[4] = {0, 0, (int32_t)W - 1, (int32_t)H - 1};
  attr("dataWindow", "box2i", box, 16);
  attr("displayWindow", "box2i", box, 16);
  const uint8_t increasingY = 0;
  attr("lineOrder", "lineOrder", &increasingY, 1);
  const float aspect = 1.0f;
  attr("pixelAspectRatio", "float", &aspect, 4);
  const float center[2] = {0.0f, 0.0f};
  attr("screenWindowCenter", "v2f", center, 8);
  attr("screenWindowWidth", "float", &aspect, 4);
  h.push_back(0);  // end of header
  const uint64_t lineBytes = 8 + (uint64_t)W * 3 * 4;
  const uint64_t table = h.size();
  for (uint32_t y = 0; y < H; ++y) {
    const uint64_t off = table + 8ull * H + (uint64_t)y * lineBytes;
    put(&off, 8);
  }
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("could not open file " + path + " for writing");
  bool ok = fwrite(h.data(), 1, h.size(), f) == h.size();
  std::vector<float> line((size_t)W * 3);
  for (uint32_t y = 0; y < H; ++y) {
    const int32_t yy = (int32_t)y, size = (int32_t)(W * 3 * 4);
    ok = ok && fwrite(&yy, 4, 1, f) == 1 && fwrite(&size, 4, 1, f) == 1;
    const float* row = rgb + (size_t)y * W * 3;
    for (uint32_t x = 0; x < W; ++x) {  // channels in B, G, R order, each a run of W floats
      line[x] = row[3 * x + 2];
      line[W + x] = row[3 * x + 1];
      line[2 * W + x] = row[3 * x];
    }
    ok = ok && fwrite(line.data(), 4, line.size(), f) == line.size();
  }
  if (fclose(f) != 0 || !ok) throw std::runtime_error("could not write file " + path);
}

}  // namespace hasty

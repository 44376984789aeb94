// Radiance RGBE (.hdr) decoding for environment maps (parseHDR, reference
// src/loaders/hdr.ts:7-93): the "#?" signature, header lines up to a blank
// line (FORMAT must be 32-bit_rle_rgbe when given), a resolution line in any
// of the eight axis orders, then flat or "new" run-length encoded
// scanlines.  Output: RGB floats, rows top to bottom, each texel
// (mantissa + 0.5) * 2^(exponent - 136) (exact in float32), zero where the
// exponent byte is 0.  Errors carry the reference's messages.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace lucent;

namespace {

struct Reader {
  const uint8_t* b;
  size_t n, pos = 0;
  // readLine: bytes up to '\n'; false at the end of the buffer.
  bool line(std::string* out) {
    const size_t start = pos;
    while (pos < n && b[pos] != 0x0a) ++pos;
    if (pos >= n) return false;
    out->assign(reinterpret_cast<const char*>(b + start), pos - start);
    ++pos;
    return true;
  }
};

std::string trim(const std::string& s) {
  // JavaScript's trim(): whitespace and line terminators at both ends.
  auto ws = [](unsigned char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; };
  size_t a = 0, e = s.size();
  while (a < e && ws(s[a])) ++a;
  while (e > a && ws(s[e - 1])) --e;
  return s.substr(a, e - a);
}

// /^([-+])([XY]) (\d+) ([-+])([XY]) (\d+)$/
bool parse_resolution(const std::string& s, char* sign1, char* axis1, uint64_t* n1, char* sign2, char* axis2,
                      uint64_t* n2) {
  size_t i = 0;
  auto part = [&](char* sign, char* axis, uint64_t* num, bool last) {
    if (i + 2 > s.size() || (s[i] != '-' && s[i] != '+') || (s[i + 1] != 'X' && s[i + 1] != 'Y')) return false;
    *sign = s[i];
    *axis = s[i + 1];
    i += 2;
    if (i >= s.size() || s[i] != ' ') return false;
    ++i;
    const size_t d0 = i;
    uint64_t v = 0;
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') {
      v = v * 10 + uint64_t(s[i] - '0');
      if (v > (1ull << 32)) v = 1ull << 32;  // far beyond any image size
      ++i;
    }
    if (i == d0) return false;
    *num = v;
    if (last) return i == s.size();
    if (i >= s.size() || s[i] != ' ') return false;
    ++i;
    return true;
  };
  return part(sign1, axis1, n1, false) && part(sign2, axis2, n2, true);
}

struct HdrError {
  const char* msg;
};

}  // namespace

extern "C" {

lucent_status lucent_decode_hdr(const uint8_t* bytes, size_t length, uint32_t* width, uint32_t* height, float* rgb,
                                size_t rgb_capacity) {
  clear_error();
  if (!bytes || !width || !height) return fail(LUCENT_E_INVALID, "null argument");
  Reader r{bytes, length};
  std::string line;
  try {
    if (!r.line(&line)) throw HdrError{"HDR: unexpected end of header"};
    if (line.compare(0, 2, "#?") != 0) throw HdrError{"HDR: missing #? signature, not a Radiance file"};
    for (;;) {
      if (!r.line(&line)) throw HdrError{"HDR: unexpected end of header"};
      line = trim(line);
      if (line.empty()) break;
      if (line.compare(0, 7, "FORMAT=") == 0 && line.size() > 7 && line.substr(7) != "32-bit_rle_rgbe") {
        static thread_local std::string msg;
        msg = "HDR: unsupported format " + line.substr(7);
        throw HdrError{msg.c_str()};
      }
    }
    if (!r.line(&line)) throw HdrError{"HDR: unexpected end of header"};
    char s1, a1, s2, a2;
    uint64_t n1, n2;
    if (!parse_resolution(trim(line), &s1, &a1, &n1, &s2, &a2, &n2) || a1 == a2)
      throw HdrError{"HDR: malformed resolution line"};
    const bool row_major_y = a1 == 'Y';
    const uint64_t w = row_major_y ? n2 : n1, h = row_major_y ? n1 : n2;
    const uint64_t scanlines = n1, scan_length = n2;
    if (w == 0 || h == 0) throw HdrError{"HDR: empty image"};
    if (w * h > (1ull << 28)) throw HdrError{"HDR: image larger than 2^28 texels"};
    *width = uint32_t(w);
    *height = uint32_t(h);
    if (!rgb) return LUCENT_OK;  // size query
    if (rgb_capacity < w * h * 3) return fail(LUCENT_E_INVALID, "rgb buffer smaller than width * height * 3 floats");
    std::memset(rgb, 0, size_t(w * h * 3) * sizeof(float));
    const bool flip_y = (row_major_y ? s1 : s2) == '+';
    const bool flip_x = (row_major_y ? s2 : s1) == '-';
    std::vector<uint8_t> rgbe(size_t(scan_length) * 4);
    const uint8_t* b = bytes;
    size_t& pos = r.pos;
    for (uint64_t s = 0; s < scanlines; ++s) {
      const bool rle = scan_length >= 8 && scan_length < 0x8000 && pos + 4 <= length && b[pos] == 2 &&
                       b[pos + 1] == 2 && (b[pos + 2] & 0x80) == 0;
      if (!rle) {
        if (pos + scan_length * 4 > length) throw HdrError{"HDR: truncated pixel data"};
        std::memcpy(rgbe.data(), b + pos, size_t(scan_length) * 4);
        pos += size_t(scan_length) * 4;
      } else {
        if (((uint32_t(b[pos + 2]) << 8) | b[pos + 3]) != scan_length) throw HdrError{"HDR: scanline width mismatch"};
        pos += 4;
        for (int c = 0; c < 4; ++c) {
          uint64_t x = 0;
          while (x < scan_length) {
            if (pos >= length) throw HdrError{"HDR: truncated run-length data"};
            uint32_t count = b[pos++];
            if (count > 128) {
              count -= 128;
              if (x + count > scan_length) throw HdrError{"HDR: run overflows scanline"};
              if (pos >= length) throw HdrError{"HDR: truncated run-length data"};
              const uint8_t v = b[pos++];
              for (uint32_t k = 0; k < count; ++k) rgbe[size_t(x++) * 4 + c] = v;
            } else {
              if (count == 0 || x + count > scan_length) throw HdrError{"HDR: bad literal run"};
              if (pos + count > length) throw HdrError{"HDR: truncated run-length data"};
              for (uint32_t k = 0; k < count; ++k) rgbe[size_t(x++) * 4 + c] = b[pos++];
            }
          }
        }
      }
      for (uint64_t i = 0; i < scan_length; ++i) {
        uint64_t x = row_major_y ? i : s, y = row_major_y ? s : i;
        if (flip_x) x = w - 1 - x;
        if (flip_y) y = h - 1 - y;
        const uint8_t e = rgbe[size_t(i) * 4 + 3];
        if (e == 0) continue;
        // Radiance's colr_color(): mantissa centred in its bucket.
        float* o = rgb + size_t(y * w + x) * 3;
        for (int c = 0; c < 3; ++c) o[c] = std::ldexp(float(rgbe[size_t(i) * 4 + c]) + 0.5f, int(e) - 136);
      }
    }
  } catch (const HdrError& err) {
    return fail(LUCENT_E_INVALID, err.msg);
  }
  return LUCENT_OK;
}

}  // extern "C"

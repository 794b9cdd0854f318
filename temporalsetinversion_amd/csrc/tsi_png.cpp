// tsi_png.cpp -- the reference's frame dump (src/main.rs:693-715): the RGBA32F frame converted
// to RGBA8 and written as a PNG.  The writer is self-contained: zlib stream of stored
// (uncompressed) deflate blocks, CRC-32 per chunk, Adler-32 per stream (RFC 1950/1951, PNG 1.2).
#include <cstdio>
#include <cstring>
#include <vector>

#include "tsi.h"

namespace {

uint32_t crc_table[256];
bool crc_ready = false;

void make_crc_table() {
    for (uint32_t n = 0; n < 256; ++n) {
        uint32_t c = n;
        for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
    crc_ready = true;
}

uint32_t crc32(const uint8_t *p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c;
}

void put32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back(uint8_t(x >> 24));
    v.push_back(uint8_t(x >> 16));
    v.push_back(uint8_t(x >> 8));
    v.push_back(uint8_t(x));
}

void chunk(std::vector<uint8_t> &out, const char type[4], const std::vector<uint8_t> &data) {
    put32(out, uint32_t(data.size()));
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put32(out, crc32(out.data() + at, data.size() + 4) ^ 0xFFFFFFFFu);
}

}  // namespace

extern "C" {

int tsi_write_png(const char *path, const uint8_t *rgba8, uint32_t width, uint32_t height) {
    if (!path || !rgba8 || width == 0 || height == 0) return TSI_E_ARG;
    if (!crc_ready) make_crc_table();
    std::vector<uint8_t> raw, z, ihdr, out;
    try {
        const size_t row = size_t(width) * 4u;
        raw.reserve((row + 1u) * height);
        for (uint32_t y = 0; y < height; ++y) {   // filter type 0 for every scanline
            raw.push_back(0);
            raw.insert(raw.end(), rgba8 + row * y, rgba8 + row * (y + 1u));
        }
        // zlib: CMF/FLG (deflate, 32K window, no dictionary, check bits), stored blocks, Adler-32
        z.push_back(0x78);
        z.push_back(0x01);
        uint32_t a = 1, b = 0;
        for (size_t i = 0; i < raw.size(); i += 65535u) {
            const size_t n = raw.size() - i < 65535u ? raw.size() - i : 65535u;
            z.push_back(i + n == raw.size() ? 1 : 0);   // BFINAL, BTYPE = 00
            z.push_back(uint8_t(n));
            z.push_back(uint8_t(n >> 8));
            z.push_back(uint8_t(~n));
            z.push_back(uint8_t(~n >> 8));
            z.insert(z.end(), raw.begin() + long(i), raw.begin() + long(i + n));
            for (size_t k = i; k < i + n; ++k) {
                a = (a + raw[k]) % 65521u;
                b = (b + a) % 65521u;
            }
        }
        put32(z, (b << 16) | a);
        put32(ihdr, width);
        put32(ihdr, height);
        ihdr.push_back(8);   // bit depth
        ihdr.push_back(6);   // colour type RGBA
        ihdr.push_back(0);   // deflate
        ihdr.push_back(0);   // adaptive filtering
        ihdr.push_back(0);   // no interlace
        static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
        out.insert(out.end(), sig, sig + 8);
        chunk(out, "IHDR", ihdr);
        chunk(out, "IDAT", z);
        chunk(out, "IEND", {});
    } catch (...) {
        return TSI_E_OOM;
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return TSI_E_ARG;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    std::fclose(f);
    return ok ? TSI_OK : TSI_E_ARG;
}

}  // extern "C"

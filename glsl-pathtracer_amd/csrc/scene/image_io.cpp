// PNG and Radiance HDR decoding (see image_io.h). PNG follows the PNG
// specification (ISO/IEC 15948): chunk stream, zlib-compressed IDAT, per-row
// filters None/Sub/Up/Average/Paeth. HDR follows the Radiance picture format:
// text header, resolution string, scanlines flat or run-length encoded per
// component.
#include "image_io.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

namespace glslpt {
namespace {

bool ReadFile(const std::string& path, std::vector<uint8_t>& data) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    data.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

uint32_t BE32(const uint8_t* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

int Paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

}  // namespace

bool LoadPNG(const std::string& path, int& width, int& height, std::vector<uint8_t>& rgba, std::string& err) {
    std::vector<uint8_t> f;
    if (!ReadFile(path, f)) { err = "cannot open " + path; return false; }
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) { err = path + ": not a PNG file"; return false; }
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    size_t pos = 8;
    bool have_ihdr = false, end = false;
    while (!end) {
        if (pos + 8 > f.size()) { err = path + ": truncated PNG"; return false; }
        const uint32_t len = BE32(&f[pos]);
        const char* type = reinterpret_cast<const char*>(&f[pos + 4]);
        if (len > f.size() || pos + 12 + len > f.size()) { err = path + ": truncated chunk"; return false; }
        const uint8_t* d = &f[pos + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) { err = path + ": bad IHDR"; return false; }
            W = BE32(d); H = BE32(d + 4); depth = d[8]; ctype = d[9]; interlace = d[12];
            have_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            end = true;
        }
        pos += 12 + len;
    }
    if (!have_ihdr || W == 0 || H == 0 || W > (1u << 15) || H > (1u << 15)) { err = path + ": bad PNG size"; return false; }
    if (interlace != 0) { err = path + ": interlaced PNG is not supported"; return false; }
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: err = path + ": unknown PNG colour type"; return false;
    }
    const bool depth_ok = depth == 8 || depth == 16 || ((ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4));
    if (!depth_ok) { err = path + ": unsupported PNG bit depth"; return false; }
    if (ctype == 3 && plte.empty()) { err = path + ": palette PNG without PLTE"; return false; }
    const size_t bits_pp = static_cast<size_t>(channels) * depth;
    const size_t row_bytes = (static_cast<size_t>(W) * bits_pp + 7) / 8;
    const size_t bpp = std::max<size_t>(1, bits_pp / 8);
    std::vector<uint8_t> raw((row_bytes + 1) * H);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) { err = "zlib init failed"; return false; }
    zs.next_in = idat.data();
    zs.avail_in = static_cast<uInt>(idat.size());
    zs.next_out = raw.data();
    zs.avail_out = static_cast<uInt>(raw.size());
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_OK && zr != Z_BUF_ERROR) || got != raw.size()) {
        err = path + ": corrupt PNG image data";
        return false;
    }
    // undo the row filters in place
    std::vector<uint8_t> img(row_bytes * H);
    for (uint32_t y = 0; y < H; ++y) {
        const uint8_t ft = raw[y * (row_bytes + 1)];
        const uint8_t* src = &raw[y * (row_bytes + 1) + 1];
        uint8_t* dst = &img[y * row_bytes];
        const uint8_t* up = y ? &img[(y - 1) * row_bytes] : nullptr;
        for (size_t i = 0; i < row_bytes; ++i) {
            const int a = i >= bpp ? dst[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v = src[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += Paeth(a, b, c); break;
                default: err = path + ": bad PNG filter"; return false;
            }
            dst[i] = static_cast<uint8_t>(v & 0xff);
        }
    }
    width = static_cast<int>(W);
    height = static_cast<int>(H);
    rgba.assign(static_cast<size_t>(W) * H * 4, 255);
    auto sample = [&](const uint8_t* row, uint32_t x, int ch) -> int {
        // channel ch of pixel x as an integer of `depth` bits
        if (depth == 8) return row[x * channels + ch];
        if (depth == 16) return (row[2 * (x * channels + ch)] << 8) | row[2 * (x * channels + ch) + 1];
        const size_t bit = static_cast<size_t>(x) * depth;
        return (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1 << depth) - 1);
    };
    auto to8 = [&](int v) -> uint8_t {
        if (depth == 16) return static_cast<uint8_t>(v >> 8);
        if (depth < 8) return static_cast<uint8_t>(v * (255 / ((1 << depth) - 1)));
        return static_cast<uint8_t>(v);
    };
    for (uint32_t y = 0; y < H; ++y) {
        const uint8_t* row = &img[y * row_bytes];
        for (uint32_t x = 0; x < W; ++x) {
            uint8_t* o = &rgba[(static_cast<size_t>(y) * W + x) * 4];
            if (ctype == 3) {
                const int idx = sample(row, x, 0);
                if (static_cast<size_t>(3 * idx + 2) < plte.size()) {
                    o[0] = plte[3 * idx]; o[1] = plte[3 * idx + 1]; o[2] = plte[3 * idx + 2];
                } else {
                    o[0] = o[1] = o[2] = 0;
                }
                o[3] = static_cast<size_t>(idx) < trns.size() ? trns[idx] : 255;
                continue;
            }
            int raw_c[4] = {0, 0, 0, 0};
            for (int c = 0; c < channels; ++c) raw_c[c] = sample(row, x, c);
            if (channels <= 2) {
                o[0] = o[1] = o[2] = to8(raw_c[0]);
                o[3] = channels == 2 ? to8(raw_c[1]) : 255;
                if (channels == 1 && trns.size() >= 2 && raw_c[0] == ((trns[0] << 8) | trns[1])) o[3] = 0;
            } else {
                o[0] = to8(raw_c[0]); o[1] = to8(raw_c[1]); o[2] = to8(raw_c[2]);
                o[3] = channels == 4 ? to8(raw_c[3]) : 255;
                if (channels == 3 && trns.size() >= 6 && raw_c[0] == ((trns[0] << 8) | trns[1]) &&
                    raw_c[1] == ((trns[2] << 8) | trns[3]) && raw_c[2] == ((trns[4] << 8) | trns[5]))
                    o[3] = 0;
            }
        }
    }
    return true;
}

bool LoadHDR(const std::string& path, int& width, int& height, std::vector<float>& rgb, std::string& err) {
    std::vector<uint8_t> f;
    if (!ReadFile(path, f)) { err = "cannot open " + path; return false; }
    size_t pos = 0;
    auto line = [&]() {
        std::string s;
        while (pos < f.size() && f[pos] != '\n') s.push_back(static_cast<char>(f[pos++]));
        if (pos < f.size()) ++pos;
        return s;
    };
    const std::string magic = line();
    if (magic != "#?RADIANCE" && magic != "#?RGBE") { err = path + ": not a Radiance HDR file"; return false; }
    bool rle_rgbe = false;
    for (;;) {
        if (pos >= f.size()) { err = path + ": truncated HDR header"; return false; }
        const std::string s = line();
        if (s.empty()) break;
        if (s == "FORMAT=32-bit_rle_rgbe") rle_rgbe = true;
    }
    if (!rle_rgbe) { err = path + ": unsupported HDR format"; return false; }
    int H = 0, W = 0;
    if (std::sscanf(line().c_str(), "-Y %d +X %d", &H, &W) != 2 || W <= 0 || H <= 0 || W > (1 << 15) || H > (1 << 15)) {
        err = path + ": unsupported HDR orientation or size";
        return false;
    }
    width = W;
    height = H;
    rgb.assign(static_cast<size_t>(W) * H * 3, 0.f);
    std::vector<uint8_t> scan(static_cast<size_t>(W) * 4);
    auto emit = [&](int y) {
        for (int x = 0; x < W; ++x) {
            const uint8_t* e = &scan[4 * x];
            float* o = &rgb[(static_cast<size_t>(y) * W + x) * 3];
            if (e[3] == 0) { o[0] = o[1] = o[2] = 0.f; continue; }
            const float s = std::ldexp(1.0f, static_cast<int>(e[3]) - 136);
            o[0] = e[0] * s; o[1] = e[1] * s; o[2] = e[2] * s;
        }
    };
    auto flat_rest = [&](int y0) {
        for (int y = y0; y < H; ++y) {
            if (pos + 4 * static_cast<size_t>(W) > f.size()) return false;
            std::memcpy(scan.data(), &f[pos], 4 * static_cast<size_t>(W));
            pos += 4 * static_cast<size_t>(W);
            emit(y);
        }
        return true;
    };
    if (W < 8 || W >= 32768) {
        if (!flat_rest(0)) { err = path + ": truncated HDR data"; return false; }
        return true;
    }
    for (int y = 0; y < H; ++y) {
        if (pos + 4 > f.size()) { err = path + ": truncated HDR data"; return false; }
        if (f[pos] != 2 || f[pos + 1] != 2 || (f[pos + 2] & 0x80)) {
            // not run-length encoded: the rest of the image is flat RGBE
            if (!flat_rest(y)) { err = path + ": truncated HDR data"; return false; }
            return true;
        }
        if (((f[pos + 2] << 8) | f[pos + 3]) != W) { err = path + ": HDR scanline width mismatch"; return false; }
        pos += 4;
        for (int c = 0; c < 4; ++c) {
            int x = 0;
            while (x < W) {
                if (pos >= f.size()) { err = path + ": truncated HDR data"; return false; }
                int n = f[pos++];
                if (n > 128) {
                    n -= 128;
                    if (pos >= f.size() || x + n > W) { err = path + ": bad HDR run"; return false; }
                    const uint8_t v = f[pos++];
                    for (int k = 0; k < n; ++k) scan[4 * (x++) + c] = v;
                } else {
                    if (n == 0 || pos + n > f.size() || x + n > W) { err = path + ": bad HDR run"; return false; }
                    for (int k = 0; k < n; ++k) scan[4 * (x++) + c] = f[pos++];
                }
            }
        }
        emit(y);
    }
    return true;
}

}  // namespace glslpt

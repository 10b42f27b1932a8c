// Image files of the reference's front end; see include/lfrv_ref_image.hpp.
#include "../../include/lfrv_ref_image.hpp"

#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>

namespace lfrv_ref {

std::vector<float> Image::planar() const
{
    const size_t n = _w * _h;
    std::vector<float> v(3 * n);
    std::copy(_r.begin(), _r.end(), v.begin());
    std::copy(_g.begin(), _g.end(), v.begin() + n);
    std::copy(_b.begin(), _b.end(), v.begin() + 2 * n);
    return v;
}

void Image::set_planar(const float *rgb)
{
    const size_t n = _w * _h;
    _r.assign(rgb, rgb + n);
    _g.assign(rgb + n, rgb + 2 * n);
    _b.assign(rgb + 2 * n, rgb + 3 * n);
}

std::string pattern_name(const std::string &pattern, int index)
{
    std::vector<char> buf(pattern.size() + 64);
    std::snprintf(buf.data(), buf.size(), pattern.c_str(), index);
    return std::string(buf.data());
}

static bool host_little_endian()
{
    const uint16_t one = 1;
    return *reinterpret_cast<const uint8_t *>(&one) == 1;
}

static void swap_bytes(float *p, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        uint8_t *b = reinterpret_cast<uint8_t *>(p + i);
        std::swap(b[0], b[3]);
        std::swap(b[1], b[2]);
    }
}

struct FileCloser {
    void operator()(FILE *f) const { if (f) std::fclose(f); }
};
using File = std::unique_ptr<FILE, FileCloser>;

// header tokens separated by whitespace, '#' comment lines skipped
static bool pfm_token(FILE *f, char *out, size_t cap)
{
    int c;
    for (;;) {
        c = std::fgetc(f);
        if (c == EOF) return false;
        if (c == '#') {
            while (c != '\n' && c != EOF) c = std::fgetc(f);
            continue;
        }
        if (!std::isspace(c)) break;
    }
    size_t k = 0;
    while (c != EOF && !std::isspace(c) && k + 1 < cap) {
        out[k++] = (char)c;
        c = std::fgetc(f);
    }
    out[k] = 0;
    return k > 0;  // the single whitespace after the scale has been consumed
}

bool load_pfm(const std::string &path, Image &img, int *channels)
{
    File f(std::fopen(path.c_str(), "rb"));
    if (!f) return false;
    char t[64], sw[64], sh[64], ss[64];
    if (!pfm_token(f.get(), t, sizeof t) || !pfm_token(f.get(), sw, sizeof sw) || !pfm_token(f.get(), sh, sizeof sh) ||
        !pfm_token(f.get(), ss, sizeof ss))
        return false;
    if (t[0] != 'P' || (t[1] != 'F' && t[1] != 'f')) return false;
    const int nc = t[1] == 'F' ? 3 : 1;
    const long W = std::atol(sw), H = std::atol(sh);
    const double scale = std::atof(ss);
    if (W <= 0 || H <= 0) return false;
    const bool file_le = scale < 0.0;
    const bool swap = file_le != host_little_endian();
    img = Image((size_t)W, (size_t)H);
    std::vector<float> row((size_t)W * nc);
    for (long y = 0; y < H; y++) {
        if (std::fread(row.data(), sizeof(float), row.size(), f.get()) != row.size()) return false;
        if (swap) swap_bytes(row.data(), row.size());
        const size_t o = (size_t)(H - 1 - y) * W;  // bottom-up rows
        for (long x = 0; x < W; x++) {
            if (nc == 3) {
                img._r[o + x] = row[3 * x];
                img._g[o + x] = row[3 * x + 1];
                img._b[o + x] = row[3 * x + 2];
            } else {
                img._r[o + x] = img._g[o + x] = img._b[o + x] = row[x];
            }
        }
    }
    if (channels) *channels = nc;
    return true;
}

static bool save_pfm_rows(const std::string &path, size_t W, size_t H, int nc, const float *const *planes)
{
    File f(std::fopen(path.c_str(), "wb"));
    if (!f) return false;
    std::fprintf(f.get(), "P%c\n%zu %zu\n%d.0\n", nc == 3 ? 'F' : 'f', W, H, host_little_endian() ? -1 : 1);
    std::vector<float> row(W * nc);
    for (size_t y = 0; y < H; y++) {
        const size_t o = (H - 1 - y) * W;
        for (size_t x = 0; x < W; x++)
            for (int c = 0; c < nc; c++) row[x * nc + c] = planes[c][o + x];
        if (std::fwrite(row.data(), sizeof(float), row.size(), f.get()) != row.size()) return false;
    }
    return true;
}

bool save_pfm(const std::string &path, const Image &img)
{
    const float *p[3] = {img._r.data(), img._g.data(), img._b.data()};
    return save_pfm_rows(path, img._w, img._h, 3, p);
}

bool save_pfm_gray(const std::string &path, size_t w, size_t h, const float *v)
{
    const float *p[1] = {v};
    return save_pfm_rows(path, w, h, 1, p);
}

// ---- PNG ---------------------------------------------------------------------

static uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

static void put32(std::vector<uint8_t> &v, uint32_t x)
{
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

static int paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

bool load_png(const std::string &path, Image &img)
{
    File f(std::fopen(path.c_str(), "rb"));
    if (!f) return false;
    std::vector<uint8_t> data;
    {
        uint8_t buf[65536];
        size_t k;
        while ((k = std::fread(buf, 1, sizeof buf, f.get())) > 0) data.insert(data.end(), buf, buf + k);
    }
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (data.size() < 8 || std::memcmp(data.data(), sig, 8) != 0) return false;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat;
    size_t pos = 8;
    while (pos + 12 <= data.size()) {
        const uint32_t len = be32(&data[pos]);
        if (pos + 12 + (size_t)len > data.size()) return false;
        const char *type = reinterpret_cast<const char *>(&data[pos + 4]);
        const uint8_t *body = &data[pos + 8];
        if (!std::strncmp(type, "IHDR", 4)) {
            W = be32(body);
            H = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (!std::strncmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::strncmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + (size_t)len;
    }
    int bpp;
    switch (ctype) {
    case 0: bpp = 1; break;  // grey
    case 2: bpp = 3; break;  // RGB
    case 4: bpp = 2; break;  // grey + alpha
    case 6: bpp = 4; break;  // RGBA
    default: return false;   // palette and others are not supported
    }
    if (depth != 8 || interlace != 0 || W == 0 || H == 0) return false;
    const size_t stride = (size_t)W * bpp;
    std::vector<uint8_t> raw((stride + 1) * H);
    uLongf rawlen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rawlen, idat.data(), (uLong)idat.size()) != Z_OK || rawlen != raw.size()) return false;
    std::vector<uint8_t> cur(stride), prev(stride, 0);
    img = Image(W, H);
    for (uint32_t y = 0; y < H; y++) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t *src = &raw[y * (stride + 1) + 1];
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            int v = src[i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: v += paeth(a, b, c); break;
            default: return false;
            }
            cur[i] = (uint8_t)v;
        }
        const size_t o = (size_t)y * W;
        for (uint32_t x = 0; x < W; x++) {
            const uint8_t *p = &cur[(size_t)x * bpp];
            // v / 255 in double, then to float, as the double planes reach the device
            float r, g, bl;
            if (bpp <= 2) r = g = bl = (float)(p[0] / 255.0);
            else {
                r = (float)(p[0] / 255.0);
                g = (float)(p[1] / 255.0);
                bl = (float)(p[2] / 255.0);
            }
            img._r[o + x] = r;
            img._g[o + x] = g;
            img._b[o + x] = bl;
        }
        std::swap(cur, prev);
    }
    return true;
}

static void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &body)
{
    put32(out, (uint32_t)body.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), body.begin(), body.end());
    const uint32_t crc = (uint32_t)crc32(0L, &out[start], (uInt)(out.size() - start));
    put32(out, crc);
}

bool save_png(const std::string &path, const Image &img)
{
    const size_t W = img._w, H = img._h, stride = 3 * W;
    if (W == 0 || H == 0) return false;
    std::vector<uint8_t> raw((stride + 1) * H);
    // clamp(int(v * 255.0), 0, 255) in double, as the reference's save of its double planes
    auto q = [](float v) {
        const double s = (double)v * 255.0;
        if (!(s > 0.0)) return (uint8_t)0;  // negative and NaN
        return (uint8_t)(s >= 255.0 ? 255 : (int)s);
    };
    for (size_t y = 0; y < H; y++) {
        uint8_t *d = &raw[y * (stride + 1)];
        *d++ = 0;  // filter: none
        for (size_t x = 0; x < W; x++) {
            const size_t o = y * W + x;
            *d++ = q(img._r[o]);
            *d++ = q(img._g[o]);
            *d++ = q(img._b[o]);
        }
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return false;
    z.resize(zlen);
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)W);
    put32(ihdr, (uint32_t)H);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, no filter method, no interlace
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    File f(std::fopen(path.c_str(), "wb"));
    return f && std::fwrite(out.data(), 1, out.size(), f.get()) == out.size();
}

static bool ends_with(const std::string &s, const char *suf)
{
    const size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; i++)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

bool load_image(const std::string &path, Image &img)
{
    if (ends_with(path, ".pfm")) return load_pfm(path, img);
    if (ends_with(path, ".png")) return load_png(path, img);
    return false;
}

bool save_image(const std::string &path, const Image &img)
{
    if (ends_with(path, ".pfm")) return save_pfm(path, img);
    if (ends_with(path, ".png")) return save_png(path, img);
    return false;
}

}  // namespace lfrv_ref

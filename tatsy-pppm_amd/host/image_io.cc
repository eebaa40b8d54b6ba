// image_io.cc -- image files (pppm_image_write / pppm_image_load and the
// containers behind pppm_image_save).
//
// Replaces Image::save / load (sources/image.cc:215-240) and its writers and
// readers: 8-bit PNG and BMP after the gamma of Image::gamma (:202-213) with
// the byte conversion of Image::toByte (:530-533), Radiance RGBE .hdr with the
// channel encoding of HDRPixel (:62-98) and the reader of Image::loadHdr
// (:296-434), and a linear FP32 dump as PFM.  The PNG stream is zlib with
// stored (uncompressed) deflate blocks: any decoder reads it, and no
// compression library is needed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pppm.h"
#include "image_io.h"

namespace {

std::string ext_of(const char* path) {
    const std::string p(path);
    const size_t dot = p.find_last_of('.');
    std::string e = dot == std::string::npos ? "" : p.substr(dot);
    for (char& c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

// ---- byte encodings (the same arithmetic on the host and in the device
// encoder of pppm_image_save) ----
inline unsigned char to_byte(double d) {  // sources/image.cc:530-533
    d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
    return (unsigned char)(255.0 * d);
}

inline void to_rgbe(const double c[3], unsigned char out[4]) {  // sources/image.cc:73-89
    double r = c[0] > 0.0 ? c[0] : 0.0, g = c[1] > 0.0 ? c[1] : 0.0, b = c[2] > 0.0 ? c[2] : 0.0;
    double d = r > g ? r : g;
    d = d > b ? d : b;
    if (!(d > 1.0e-32)) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e;
    const double m = std::frexp(d, &e);
    d = m * 256.0 / d;
    out[0] = (unsigned char)(r * d);
    out[1] = (unsigned char)(g * d);
    out[2] = (unsigned char)(b * d);
    out[3] = (unsigned char)(e + 128);
}

// ---- containers ----
uint32_t crc_table[256];
bool crc_ready = false;
uint32_t crc32(const unsigned char* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    if (!crc_ready) {
        for (uint32_t k = 0; k < 256; k++) {
            uint32_t v = k;
            for (int j = 0; j < 8; j++) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            crc_table[k] = v;
        }
        crc_ready = true;
    }
    for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<unsigned char>& v, uint32_t x) {
    v.push_back((unsigned char)(x >> 24));
    v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8));
    v.push_back((unsigned char)x);
}

void png_chunk(std::FILE* fp, const char* type, const std::vector<unsigned char>& data) {
    std::vector<unsigned char> buf;
    put_be32(buf, (uint32_t)data.size());
    buf.insert(buf.end(), type, type + 4);
    buf.insert(buf.end(), data.begin(), data.end());
    const uint32_t c = crc32(buf.data() + 4, buf.size() - 4) ^ 0xFFFFFFFFu;
    put_be32(buf, c);
    std::fwrite(buf.data(), 1, buf.size(), fp);
}

}  // namespace

// Writers of encoded pixels, shared with the device encoder (capi.hip).
// rgb8: [h][w][3] top row first; rgbe: [h][w][4]; f32: [h][w][3] top row first.
int image_write_png8(const char* path, const unsigned char* rgb8, int32_t w, int32_t h) {
    std::FILE* fp = std::fopen(path, "wb");
    if (!fp) return PPPM_ERR_INVALID;
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::fwrite(sig, 1, 8, fp);
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)w);
    put_be32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, filter set 0, no interlace
    png_chunk(fp, "IHDR", ihdr);
    // raw scanlines (filter byte 0), zlib-wrapped in stored deflate blocks
    const size_t row = 1 + 3 * (size_t)w;
    std::vector<unsigned char> raw(row * (size_t)h);
    for (int32_t y = 0; y < h; y++) {
        raw[row * y] = 0;
        std::memcpy(&raw[row * y + 1], rgb8 + 3 * (size_t)w * y, 3 * (size_t)w);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        z.push_back(pos + len == raw.size() ? 1 : 0);
        z.push_back((unsigned char)(len & 0xFF));
        z.push_back((unsigned char)(len >> 8));
        z.push_back((unsigned char)(~len & 0xFF));
        z.push_back((unsigned char)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (long)pos, raw.begin() + (long)(pos + len));
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;  // Adler-32
    for (unsigned char c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    png_chunk(fp, "IDAT", z);
    png_chunk(fp, "IEND", {});
    const bool ok = std::ferror(fp) == 0;
    return std::fclose(fp) == 0 && ok ? PPPM_OK : PPPM_ERR_INVALID;
}

int image_write_bmp8(const char* path, const unsigned char* rgb8, int32_t w, int32_t h) {
    std::FILE* fp = std::fopen(path, "wb");
    if (!fp) return PPPM_ERR_INVALID;
    const uint32_t row = (3u * (uint32_t)w + 3u) & ~3u;
    const uint32_t size = 54 + row * (uint32_t)h;
    unsigned char hdr[54] = {'B', 'M'};
    auto le32 = [&](int at, uint32_t x) {
        for (int k = 0; k < 4; k++) hdr[at + k] = (unsigned char)(x >> (8 * k));
    };
    le32(2, size);
    le32(10, 54);
    le32(14, 40);
    le32(18, (uint32_t)w);
    le32(22, (uint32_t)h);  // positive height: rows bottom-up
    hdr[26] = 1;
    hdr[28] = 24;
    le32(34, row * (uint32_t)h);
    std::fwrite(hdr, 1, 54, fp);
    std::vector<unsigned char> line(row, 0);
    for (int32_t y = h - 1; y >= 0; y--) {
        for (int32_t x = 0; x < w; x++) {
            const unsigned char* p = rgb8 + 3 * ((size_t)w * y + x);
            line[3 * x] = p[2];
            line[3 * x + 1] = p[1];
            line[3 * x + 2] = p[0];
        }
        std::fwrite(line.data(), 1, row, fp);
    }
    const bool ok = std::ferror(fp) == 0;
    return std::fclose(fp) == 0 && ok ? PPPM_OK : PPPM_ERR_INVALID;
}

int image_write_rgbe(const char* path, const unsigned char* rgbe, int32_t w, int32_t h) {
    std::FILE* fp = std::fopen(path, "wb");
    if (!fp) return PPPM_ERR_INVALID;
    std::fprintf(fp, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y %d +X %d\n", (int)h, (int)w);
    std::vector<unsigned char> out;
    for (int32_t y = 0; y < h; y++) {
        const unsigned char* line = rgbe + 4 * (size_t)w * y;
        if (w < 8 || w > 0x7FFF) {  // the run-length scheme covers widths 8 .. 32767
            out.insert(out.end(), line, line + 4 * (size_t)w);
            continue;
        }
        out.push_back(2);
        out.push_back(2);
        out.push_back((unsigned char)(w >> 8));
        out.push_back((unsigned char)(w & 0xFF));
        for (int c = 0; c < 4; c++) {
            // runs of >= 3 equal bytes as (128 + n, byte), the rest as literal dumps of <= 128
            int32_t x = 0;
            while (x < w) {
                int32_t run = 1;
                while (x + run < w && run < 127 && line[4 * (x + run) + c] == line[4 * x + c]) run++;
                if (run >= 3) {
                    out.push_back((unsigned char)(128 + run));
                    out.push_back(line[4 * x + c]);
                    x += run;
                    continue;
                }
                int32_t n = 0;
                while (x + n < w && n < 128) {
                    int32_t r = 1;
                    while (x + n + r < w && r < 3 && line[4 * (x + n + r) + c] == line[4 * (x + n) + c]) r++;
                    if (r >= 3) break;
                    n++;
                }
                out.push_back((unsigned char)n);
                for (int32_t k = 0; k < n; k++) out.push_back(line[4 * (x + k) + c]);
                x += n;
            }
        }
    }
    std::fwrite(out.data(), 1, out.size(), fp);
    const bool ok = std::ferror(fp) == 0;
    return std::fclose(fp) == 0 && ok ? PPPM_OK : PPPM_ERR_INVALID;
}

int image_write_pfm(const char* path, const float* rgb, int32_t w, int32_t h) {
    std::FILE* fp = std::fopen(path, "wb");
    if (!fp) return PPPM_ERR_INVALID;
    std::fprintf(fp, "PF\n%d %d\n-1.0\n", (int)w, (int)h);  // negative scale: little endian
    for (int32_t y = h - 1; y >= 0; y--)                      // PFM rows run bottom to top
        std::fwrite(rgb + 3 * (size_t)w * y, sizeof(float), 3 * (size_t)w, fp);
    const bool ok = std::ferror(fp) == 0;
    return std::fclose(fp) == 0 && ok ? PPPM_OK : PPPM_ERR_INVALID;
}

namespace {

bool read_all(const char* path, std::vector<unsigned char>* data) {
    std::FILE* fp = std::fopen(path, "rb");
    if (!fp) return false;
    std::fseek(fp, 0, SEEK_END);
    const long n = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (n < 0) return std::fclose(fp), false;
    data->resize((size_t)n);
    const size_t got = n ? std::fread(data->data(), 1, (size_t)n, fp) : 0;
    std::fclose(fp);
    return got == (size_t)n;
}

// a header line ending in '\n' at *pos
bool next_line(const std::vector<unsigned char>& d, size_t* pos, std::string* line) {
    line->clear();
    while (*pos < d.size() && d[*pos] != '\n') line->push_back((char)d[(*pos)++]);
    if (*pos >= d.size()) return false;
    (*pos)++;
    return true;
}

// Radiance .hdr: "#?RADIANCE" (or "#?RGBE"), header lines to a blank line,
// "-Y h +X w", then scanlines in the run-length scheme (what the reference
// reads, sources/image.cc:296-434) or flat RGBE quads.  Decoded as the
// reference does: c * 2^(e - 128) / 256.
bool load_hdr(const char* path, std::vector<double>* rgb, int32_t* w, int32_t* h, std::string* err) {
    std::vector<unsigned char> d;
    if (!read_all(path, &d)) return *err = std::string("cannot read ") + path, false;
    size_t pos = 0;
    std::string line;
    if (!next_line(d, &pos, &line) || (line != "#?RADIANCE" && line != "#?RGBE"))
        return *err = "not a Radiance HDR file", false;
    bool rgbe = true;
    while (next_line(d, &pos, &line) && !line.empty())
        if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe") rgbe = false;
    if (!rgbe) return *err = "only FORMAT=32-bit_rle_rgbe is read", false;
    if (!next_line(d, &pos, &line)) return *err = "missing resolution line", false;
    char ay[8], ax[8];
    int hh = 0, ww = 0;
    if (std::sscanf(line.c_str(), "%7s %d %7s %d", ay, &hh, ax, &ww) != 4 || std::strcmp(ay, "-Y") != 0 ||
        std::strcmp(ax, "+X") != 0 || ww <= 0 || hh <= 0)
        return *err = "resolution line must be '-Y h +X w': " + line, false;
    std::vector<unsigned char> px(4 * (size_t)ww * hh);
    for (int y = 0; y < hh; y++) {
        unsigned char* out = &px[4 * (size_t)ww * y];
        if (pos + 4 > d.size()) return *err = "truncated scanline", false;
        const bool rle = ww >= 8 && ww <= 0x7FFF && d[pos] == 2 && d[pos + 1] == 2 && !(d[pos + 2] & 0x80);
        if (!rle) {
            if (pos + 4 * (size_t)ww > d.size()) return *err = "truncated flat scanline", false;
            std::memcpy(out, &d[pos], 4 * (size_t)ww);
            pos += 4 * (size_t)ww;
            continue;
        }
        if (((int)d[pos + 2] << 8 | d[pos + 3]) != ww) return *err = "scanline width mismatch", false;
        pos += 4;
        for (int c = 0; c < 4; c++) {
            int x = 0;
            while (x < ww) {
                if (pos >= d.size()) return *err = "truncated run", false;
                const int info = d[pos++];
                if (info > 128) {
                    const int n = info - 128;
                    if (x + n > ww || pos >= d.size()) return *err = "bad run", false;
                    for (int k = 0; k < n; k++) out[4 * (x + k) + c] = d[pos];
                    pos++;
                    x += n;
                } else {
                    if (info == 0 || x + info > ww || pos + info > d.size()) return *err = "bad literal dump", false;
                    for (int k = 0; k < info; k++) out[4 * (x + k) + c] = d[pos++];
                    x += info;
                }
            }
        }
    }
    rgb->resize(3 * (size_t)ww * hh);
    for (size_t i = 0; i < (size_t)ww * hh; i++) {
        const double s = std::pow(2.0, px[4 * i + 3] - 128.0) / 256.0;
        for (int c = 0; c < 3; c++) (*rgb)[3 * i + c] = px[4 * i + c] * s;
    }
    *w = ww;
    *h = hh;
    return true;
}

bool load_pfm(const char* path, std::vector<double>* rgb, int32_t* w, int32_t* h, std::string* err) {
    std::vector<unsigned char> d;
    if (!read_all(path, &d)) return *err = std::string("cannot read ") + path, false;
    size_t pos = 0;
    std::string l1, l2, l3;
    if (!next_line(d, &pos, &l1) || (l1 != "PF" && l1 != "Pf")) return *err = "not a PFM file", false;
    if (!next_line(d, &pos, &l2) || !next_line(d, &pos, &l3)) return *err = "truncated PFM header", false;
    int ww = 0, hh = 0;
    if (std::sscanf(l2.c_str(), "%d %d", &ww, &hh) != 2 || ww <= 0 || hh <= 0) return *err = "bad PFM size", false;
    const double scale = std::atof(l3.c_str());
    const int ch = l1 == "PF" ? 3 : 1;
    if (pos + sizeof(float) * ch * (size_t)ww * hh > d.size()) return *err = "truncated PFM data", false;
    rgb->resize(3 * (size_t)ww * hh);
    for (int y = 0; y < hh; y++)
        for (int x = 0; x < ww; x++)
            for (int c = 0; c < 3; c++) {
                unsigned char b[4];
                std::memcpy(b, &d[pos + 4 * ((size_t)ch * ((size_t)(hh - 1 - y) * ww + x) + (ch == 3 ? c : 0))], 4);
                if (scale > 0.0) std::swap(b[0], b[3]), std::swap(b[1], b[2]);  // big endian
                float f;
                std::memcpy(&f, b, 4);
                (*rgb)[3 * ((size_t)y * ww + x) + c] = f;
            }
    *w = ww;
    *h = hh;
    return true;
}

// 24-bit uncompressed BMP (what Image::saveBmp writes), as c / 255 (Image::toReal)
bool load_bmp(const char* path, std::vector<double>* rgb, int32_t* w, int32_t* h, std::string* err) {
    std::vector<unsigned char> d;
    if (!read_all(path, &d) || d.size() < 54 || d[0] != 'B' || d[1] != 'M') return *err = "not a BMP file", false;
    auto le32 = [&](size_t at) { return (uint32_t)d[at] | (uint32_t)d[at + 1] << 8 | (uint32_t)d[at + 2] << 16 | (uint32_t)d[at + 3] << 24; };
    const uint32_t off = le32(10);
    const int32_t ww = (int32_t)le32(18), hraw = (int32_t)le32(22);
    const int bits = d[28] | d[29] << 8;
    if (bits != 24 || le32(30) != 0 || ww <= 0 || hraw == 0) return *err = "only 24-bit uncompressed BMP is read", false;
    const int32_t hh = hraw < 0 ? -hraw : hraw;
    const size_t row = (3 * (size_t)ww + 3) & ~(size_t)3;
    if (off + row * hh > d.size()) return *err = "truncated BMP", false;
    rgb->resize(3 * (size_t)ww * hh);
    for (int32_t r = 0; r < hh; r++) {
        const int32_t y = hraw > 0 ? hh - 1 - r : r;  // positive height: bottom-up rows
        for (int32_t x = 0; x < ww; x++) {
            const unsigned char* p = &d[off + row * r + 3 * x];
            for (int c = 0; c < 3; c++) (*rgb)[3 * ((size_t)y * ww + x) + c] = p[2 - c] / 255.0;
        }
    }
    *w = ww;
    *h = hh;
    return true;
}

}  // namespace

extern "C" {

int pppm_image_write(const char* path, const double* rgb, int32_t width, int32_t height, double gamma) {
    if (!path || !rgb || width <= 0 || height <= 0) return PPPM_ERR_INVALID;
    const std::string e = ext_of(path);
    const size_t n = (size_t)width * height;
    if (e == ".png" || e == ".bmp") {
        const double g = gamma > 0.0 ? 1.0 / gamma : 1.0;
        std::vector<unsigned char> b(3 * n);
        for (size_t i = 0; i < 3 * n; i++) b[i] = to_byte(gamma > 0.0 ? std::pow(rgb[i], g) : rgb[i]);
        return e == ".png" ? image_write_png8(path, b.data(), width, height) : image_write_bmp8(path, b.data(), width, height);
    }
    if (e == ".hdr") {
        std::vector<unsigned char> b(4 * n);
        for (size_t i = 0; i < n; i++) to_rgbe(rgb + 3 * i, &b[4 * i]);
        return image_write_rgbe(path, b.data(), width, height);
    }
    if (e == ".pfm") {
        std::vector<float> f(3 * n);
        for (size_t i = 0; i < 3 * n; i++) f[i] = (float)rgb[i];
        return image_write_pfm(path, f.data(), width, height);
    }
    return PPPM_ERR_INVALID;
}

int pppm_image_load(const char* path, double** rgb, int32_t* width, int32_t* height, char* err, int32_t err_len) {
    if (!path || !rgb || !width || !height) return PPPM_ERR_INVALID;
    *rgb = nullptr;
    std::vector<double> v;
    std::string e;
    int32_t w = 0, h = 0;
    const std::string x = ext_of(path);
    bool ok;
    if (x == ".hdr") ok = load_hdr(path, &v, &w, &h, &e);
    else if (x == ".pfm") ok = load_pfm(path, &v, &w, &h, &e);
    else if (x == ".bmp") ok = load_bmp(path, &v, &w, &h, &e);
    else ok = false, e = "images are read from .hdr, .pfm or .bmp";
    if (!ok) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", e.c_str());
        return PPPM_ERR_INVALID;
    }
    *rgb = static_cast<double*>(std::malloc(sizeof(double) * v.size()));
    if (!*rgb) return PPPM_ERR_NOMEM;
    std::memcpy(*rgb, v.data(), sizeof(double) * v.size());
    *width = w;
    *height = h;
    return PPPM_OK;
}

void pppm_image_free(double* rgb) { std::free(rgb); }

}  // extern "C"

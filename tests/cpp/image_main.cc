// image_main.cc -- the Image class of the C++ interface (sources/image.h):
// gamma, tonemap, save / load by extension.  Writes the files it checks into
// the directory given as argv[1]; exits 0 and prints OK when every check holds.
#include <cmath>
#include <cstdio>
#include <string>

#include "../../tatsy-pppm_amd/host/pppm_amd.h"

using namespace pppm_amd;

static int failures = 0;
static void expect(bool ok, const char* what) {
    std::printf("%s: %s\n", ok ? "ok" : "FAIL", what);
    if (!ok) failures++;
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    Image img(7, 5);
    for (int y = 0; y < 5; y++)
        for (int x = 0; x < 7; x++) img.setPixel(x, y, Vector3D(0.1 * x, 0.05 * y + 0.01, 0.5));

    // gamma: every channel to the power gam (inv: 1 / gam)
    Image g = img;
    g.gamma(2.2);
    expect(std::fabs(g(3, 2).x() - std::pow(0.3, 2.2)) < 1e-15, "gamma raises to gam");
    g.gamma(2.2, true);
    expect(std::fabs(g(3, 2).x() - 0.3) < 1e-12, "inverse gamma undoes it");

    // tonemap: a uniform image maps to s (1 + s / Lw^2) / (1 + s) with s = 0.18 c / c
    Image u(4, 4);
    u.fill(Vector3D(2.0, 2.0, 2.0));
    u.tonemap();
    const double lum = 2.0, s = 0.18 * 2.0 / std::exp(std::log(lum + 1e-8));
    const double want = s * (1.0 + s / (lum * lum)) / (1.0 + s);
    expect(std::fabs(u(1, 1).x() - want) < 1e-12, "tonemap of a uniform image");

    // save / load round trips
    const std::string hdr = dir + "/img.hdr", bmp = dir + "/img.bmp", png = dir + "/img.png";
    img.save(hdr);
    Image back;
    back.loadHdr(hdr);
    // RGBE: one exponent per pixel, 8-bit mantissas relative to the largest channel
    bool close = back.width() == 7 && back.height() == 5;
    for (int y = 0; close && y < 5; y++)
        for (int x = 0; x < 7; x++) {
            const Vector3D a = img(x, y), b = back(x, y);
            const double m = std::fmax(a.x(), std::fmax(a.y(), a.z()));
            close = close && std::fabs(a.x() - b.x()) <= m / 128 && std::fabs(a.y() - b.y()) <= m / 128 &&
                    std::fabs(a.z() - b.z()) <= m / 128;
        }
    expect(close, "RGBE round trip within the 8-bit mantissa");
    img.saveBmp(bmp);
    back.load(bmp);
    expect(back.width() == 7 && std::fabs(back(6, 4).x() - std::floor(0.6 * 255.0) / 255.0) < 1e-12,
           "BMP stores truncated bytes");
    img.save(png);
    bool threw = false;
    try {
        img.saveBmp(png);
    } catch (const Error&) {
        threw = true;
    }
    expect(threw, "saveBmp refuses another extension");
    threw = false;
    try {
        back.load(dir + "/missing.hdr");
    } catch (const Error&) {
        threw = true;
    }
    expect(threw, "a missing file throws");
    std::printf(failures ? "FAILED\n" : "OK\n");
    return failures ? 1 : 0;
}

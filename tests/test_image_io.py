"""Image files (pppm_image_write / pppm_image_load): CPU only.

* Against the compiled reference's Image (when its build is present): the
  decoded pixels of our PNG / BMP equal those of the reference's writers for
  the same FP64 image (gamma and byte truncation included), our .hdr files
  decode identically through the reference's loader and ours, and the
  reference's .hdr files decode identically through both.
* Without it: exact PFM round trips, RGBE within its 8-bit mantissa, 8-bit
  formats equal floor(255 clamp(c^(1/gamma))), run-length scanlines with long
  runs, flat scanlines, and bad files rejected.
"""
import math
import os
import struct
import zlib

import numpy as np
import pytest

import oracle as O
import pppm_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def decode_png(path):
    """8-bit RGB PNG pixels [h][w][3] (all five row filters)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) & 0xFFFFFFFF == crc
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    bpp, stride = 3, 3 * w
    out = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)], np.uint8).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            cur[i] = (line[i] + p) & 0xFF
        out[y] = cur
        prev = cur
    return out.reshape(h, w, 3).astype(np.uint8)


def image(w=37, h=23, seed=1):
    rng = np.random.default_rng(seed)
    img = rng.exponential(0.4, size=(h, w, 3))
    img[0, :5] = 0.0                 # black pixels
    img[1, :] = [0.25, 0.5, 1.5]     # a constant row: long runs in the RGBE scheme
    img[2, 3] = [1e-40, 0.0, 0.0]    # below the RGBE floor
    img[3, 4] = [4000.0, 2.0, 1e-3]  # wide dynamic range in one pixel
    return img


def to_bytes(img, gamma):
    g = np.power(img, 1.0 / gamma) if gamma > 0 else img
    return (255.0 * np.clip(g, 0.0, 1.0)).astype(np.uint8)


@pytest.mark.parametrize("gamma", [2.2, 0.0])
def test_png_and_bmp_bytes(tmp_path, gamma):
    img = image()
    P.image_write(tmp_path / "a.png", img, gamma)
    np.testing.assert_array_equal(decode_png(tmp_path / "a.png"), to_bytes(img, gamma))
    P.image_write(tmp_path / "a.bmp", img, gamma)
    np.testing.assert_array_equal(P.image_load(tmp_path / "a.bmp"), to_bytes(img, gamma) / 255.0)


def test_pfm_round_trip_exact(tmp_path):
    img = image()
    P.image_write(tmp_path / "a.pfm", img)
    np.testing.assert_array_equal(P.image_load(tmp_path / "a.pfm"), img.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("w", [37, 5, 300])  # run-length scanlines, flat scanlines (w < 8), runs > 127
def test_hdr_round_trip(tmp_path, w):
    img = image(w=w)
    P.image_write(tmp_path / "a.hdr", img)
    got = P.image_load(tmp_path / "a.hdr")
    assert got.shape == img.shape
    m = img.max(axis=2, keepdims=True)
    # an 8-bit mantissa shared by the three channels: error below 2^-8 of the largest
    assert np.all(np.abs(got - img) <= m * 2.0 ** -7 + 1e-38)
    assert np.all(got[0, :5] == 0.0) and np.all(got[2, 3] == 0.0)


def test_bad_image_files_are_rejected(tmp_path):
    p = tmp_path / "x.hdr"
    p.write_bytes(b"#?NOTHDR\n\n-Y 2 +X 2\n")
    with pytest.raises(P.PPPMError):
        P.image_load(p)
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 9\n\x02\x02\x00\x09\x85")
    with pytest.raises(P.PPPMError):
        P.image_load(p)  # truncated run
    with pytest.raises(P.PPPMError):
        P.image_load(tmp_path / "missing.pfm")
    with pytest.raises(P.PPPMError):
        P.image_write(tmp_path / "x.jpg", image())


# ---- against the compiled reference -----------------------------------------
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference absent")


def ref_save(path, img):
    img = np.ascontiguousarray(img, np.float64)
    O.ref().ref_image_save(str(path).encode(), O._ptr(img), img.shape[1], img.shape[0])


def ref_load(path, cap=1 << 20):
    out = np.zeros((cap, 3), np.float64)
    w, h = O.C.c_int32(), O.C.c_int32()
    n = O.ref().ref_image_load(str(path).encode(), O._ptr(out), cap, O.C.byref(w), O.C.byref(h))
    assert n >= 0
    return out[:n].reshape(h.value, w.value, 3)


@needs_ref
def test_png_bmp_pixels_equal_reference_writers(tmp_path):
    img = image()
    # the reference applies gamma in the render loop (Image::gamma(2.2, true)),
    # then saves: give it the gamma-corrected image, ours the linear one
    ref_save(tmp_path / "r.png", np.power(img, 1.0 / 2.2))
    P.image_write(tmp_path / "o.png", img, 2.2)
    np.testing.assert_array_equal(decode_png(tmp_path / "o.png"), decode_png(tmp_path / "r.png"))
    ref_save(tmp_path / "r.bmp", np.power(img, 1.0 / 2.2))
    P.image_write(tmp_path / "o.bmp", img, 2.2)
    np.testing.assert_array_equal(P.image_load(tmp_path / "o.bmp"), P.image_load(tmp_path / "r.bmp"))
    np.testing.assert_array_equal(ref_load(tmp_path / "o.bmp"), ref_load(tmp_path / "r.bmp"))


@needs_ref
def test_hdr_files_decode_alike_in_both_loaders(tmp_path):
    img = image(w=64, h=16)
    img[5, :] = 0.7  # a run the reference's writer stores as literals and ours as a run
    P.image_write(tmp_path / "o.hdr", img)
    ref_save(tmp_path / "r.hdr", img)
    ours_by_ref, ours_by_us = ref_load(tmp_path / "o.hdr"), P.image_load(tmp_path / "o.hdr")
    ref_by_ref, ref_by_us = ref_load(tmp_path / "r.hdr"), P.image_load(tmp_path / "r.hdr")
    np.testing.assert_array_equal(ours_by_ref, ours_by_us)
    np.testing.assert_array_equal(ref_by_ref, ref_by_us)
    # the same RGBE bytes: HDRPixel's encoding, so the decoded images agree too
    np.testing.assert_array_equal(ours_by_us, ref_by_us)


def test_cpp_image_class(tmp_path):
    """Image::gamma / tonemap / save / load of the C++ interface (tests/cpp/image_main.cc)."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "image_main")
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0 and "OK" in r.stdout

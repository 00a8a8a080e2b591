"""The grayscale baseline JPEG writer of the vorticity pictures (csrc/jpeg_gray.hpp).

The pictures are compiled through a small host program and decoded with an
independent decoder (Pillow's libjpeg): the size and mode must be the ones
written and the pixels within the rounding of the DCT at quality 100 (all
quantisation steps 1), including sizes that are not multiples of 8 and the
extreme values 0 and 255.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MAIN = r"""
#include <cstdio>
#include <vector>
#include "jpeg_gray.hpp"
int main(int argc, char** argv) {
    if (argc != 6) return 2;
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]), q = std::atoi(argv[5]);
    std::vector<uint8_t> pict((size_t)w * h);
    FILE* fp = std::fopen(argv[1], "rb");
    if (!fp || std::fread(pict.data(), 1, pict.size(), fp) != pict.size()) return 3;
    std::fclose(fp);
    return tdlb::jpeg::write_gray(argv[2], pict.data(), w, h, q) ? 0 : 4;
}
"""


@pytest.fixture(scope="module")
def encoder(tmp_path_factory):
    d = tmp_path_factory.mktemp("jpeg")
    src = d / "enc.cpp"
    src.write_text(MAIN)
    exe = d / "enc"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "tdlbcpp_amd", "csrc"), "-o", str(exe),
                    str(src)], check=True)

    def encode(pict, quality=100):
        h, w = pict.shape
        raw, out = d / "in.raw", d / f"out_{w}x{h}_{quality}.jpeg"
        np.ascontiguousarray(pict, dtype=np.uint8).tofile(raw)
        subprocess.run([str(exe), str(raw), str(out), str(w), str(h), str(quality)], check=True)
        return out
    return encode


def decode(path):
    from PIL import Image
    with Image.open(path) as im:
        assert im.format == "JPEG" and im.mode == "L"
        return np.asarray(im, dtype=np.int32), im.size


def pictures():
    rng = np.random.default_rng(7)
    yy, xx = np.mgrid[0:60, 0:60]
    smooth = (127.5 + 127.5 * np.sin(xx / 7.0) * np.cos(yy / 11.0)).astype(np.uint8)
    edges = np.where((xx - 30) ** 2 + (yy - 25) ** 2 < 200, 255, 0).astype(np.uint8)
    return {
        "smooth 60x60": smooth,
        "disc with hard edges": edges,
        "noise 37x23": rng.integers(0, 256, size=(23, 37)).astype(np.uint8),
        "ramp 7x513": np.tile((np.arange(513) % 256).astype(np.uint8), (7, 1)),
        "single pixel": np.array([[200]], dtype=np.uint8),
        "all zero 8x8": np.zeros((8, 8), dtype=np.uint8),
        "all 255 16x9": np.full((9, 16), 255, dtype=np.uint8),
    }


@pytest.mark.parametrize("name", list(pictures()))
def test_quality_100_round_trips(encoder, name):
    pict = pictures()[name]
    got, size = decode(encoder(pict))
    assert size == (pict.shape[1], pict.shape[0])
    err = np.abs(got - pict.astype(np.int32))
    # all-ones quantisation: only the rounding of the coefficients and of the decoder's IDCT
    assert err.max() <= 2, err.max()  # measured: at most 1
    assert err.mean() <= 0.2


def test_lower_quality_quantises_and_still_decodes(encoder):
    pict = pictures()["smooth 60x60"]
    q100, q50 = encoder(pict, 100), encoder(pict, 50)
    got, _ = decode(q50)
    assert os.path.getsize(q50) < os.path.getsize(q100)
    mse = np.mean((got - pict.astype(np.int32)) ** 2)
    assert 10 * np.log10(255.0 ** 2 / mse) > 30.0  # PSNR of a smooth picture at quality 50

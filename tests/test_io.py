"""Host-side formats around the chain, without a GPU: the per-level TGA files of
writeMipmapsTga (include/mipmap_storage.hpp:441-479), the TGA reader, and the
oracle's premultiplied-alpha staging (include/scoped_image.hpp:233-252)."""
import os

import numpy as np
import pytest

from oracle import orc
from vk_compute_mipmaps_amd import pyramid as P


def _chain(w, h, seed):
    offs, total = P.packed_layout(w, h)
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, total, dtype=np.uint8), offs


@pytest.mark.parametrize("w,h", [(37, 5), (64, 64), (1, 9)])
def test_tga_round_trip_every_level(w, h, tmp_path):
    chain, offs = _chain(w, h, w * 1000 + h)
    base = str(tmp_path / "mips.tga")
    P.write_tga(chain, w, h, base)
    for level, off in enumerate(offs):
        lw, lh = P.level_dims(w, h, level)
        name = P.tga_level_filename(base, level)
        assert os.path.exists(name), name
        img = P.read_tga(name)
        assert img.shape == (lh, lw, 4)
        assert np.array_equal(img.reshape(-1), chain[off:off + lw * lh * 4])
        raw = open(name, "rb").read(18)
        assert (raw[2], raw[16], raw[17]) == (2, 32, 0x28)  # uncompressed, 32 bpp, top-left origin


def test_tga_level_names_follow_the_reference(tmp_path):
    assert P.tga_level_filename("a/out.tga", 0) == "a/out.tga"
    assert P.tga_level_filename("a/out.tga", 3) == "a/out.3.tga"
    assert P.tga_level_filename("a/out.x.tga", 1) == "a/out.x.1.tga"
    assert P.tga_level_filename("noext", 2) == "noext2"
    chain, _ = _chain(4, 2, 1)
    P.write_tga(chain, 4, 2, str(tmp_path / "noext"))
    assert sorted(os.listdir(tmp_path)) == ["noext", "noext1", "noext2"]


def _rle_tga(rgba, bottom_up):
    """An RLE-compressed 32-bit TGA of rgba (h, w, 4), as stb_image_write emits."""
    h, w, _ = rgba.shape
    hdr = bytes([0, 0, 10, 0, 0, 0, 0, 0, 0, 0, 0, 0, w & 255, w >> 8, h & 255, h >> 8, 32, 8 | (0 if bottom_up else 32)])
    rows = rgba[::-1] if bottom_up else rgba
    out = bytearray(hdr)
    for row in rows:
        bgra = row[:, [2, 1, 0, 3]]
        x = 0
        while x < w:  # runs of equal texels, literals otherwise, at most 128 each
            run = 1
            while x + run < w and run < 128 and np.array_equal(bgra[x + run], bgra[x]):
                run += 1
            if run > 1:
                out += bytes([0x80 | (run - 1)]) + bytes(bgra[x])
                x += run
            else:
                lit = 1
                while x + lit < w and lit < 128 and not np.array_equal(bgra[x + lit], bgra[x + lit - 1]):
                    lit += 1
                out += bytes([lit - 1]) + bgra[x:x + lit].tobytes()
                x += lit
    return bytes(out)


@pytest.mark.parametrize("bottom_up", [True, False])
def test_tga_reader_takes_run_length_files(bottom_up, tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 4, (7, 33, 4), dtype=np.uint8)  # few values: runs and literals
    img[2, 5:30] = [9, 8, 7, 6]
    name = str(tmp_path / "rle.tga")
    open(name, "wb").write(_rle_tga(img, bottom_up))
    assert np.array_equal(P.read_tga(name), img)


def test_tga_reader_takes_24_bit_bottom_up(tmp_path):
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, (3, 5, 3), dtype=np.uint8)
    hdr = bytes([0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 5, 0, 3, 0, 24, 0])
    name = str(tmp_path / "rgb.tga")
    open(name, "wb").write(hdr + img[::-1][:, :, [2, 1, 0]].tobytes())
    got = P.read_tga(name)
    assert np.array_equal(got[..., :3], img) and (got[..., 3] == 255).all()


def test_tga_errors():
    with pytest.raises(P.NvproPyramidError):
        P.read_tga("/nonexistent/file.tga")
    with pytest.raises(P.NvproPyramidError):
        P.write_tga(np.zeros(16, np.uint8), 70000, 1, "/tmp/x.tga")


def _all_pairs():
    """Every (code, alpha) pair once: texel (code, code, code, alpha)."""
    c, a = np.meshgrid(np.arange(256, dtype=np.uint8), np.arange(256, dtype=np.uint8))
    c16 = c.astype(np.uint16)
    return np.stack([c16, (c16 + 85) % 256, (c16 + 170) % 256, a], axis=-1).astype(np.uint8).reshape(-1)


def test_premultiply_oracle_pins():
    t = _all_pairs()
    exact = orc.premultiply_srgba8(t, exact=True).reshape(-1, 4)
    src = t.reshape(-1, 4)
    # alpha 255 is the identity (encode(decode(c)) == c), alpha 0 gives black,
    # alpha bytes are kept, and the result is monotone in alpha.
    assert np.array_equal(exact[src[:, 3] == 255], src[src[:, 3] == 255])
    assert (exact[src[:, 3] == 0, :3] == 0).all()
    assert np.array_equal(exact[:, 3], src[:, 3])
    grid = exact[:, 0].reshape(256, 256)  # [alpha][code]
    assert (np.diff(grid.astype(np.int16), axis=0) >= 0).all()
    # Within one code of the reference's FP32 powf formulas everywhere.
    ref = orc.premultiply_srgba8(t, exact=False).reshape(-1, 4)
    d = np.abs(exact.astype(np.int16) - ref.astype(np.int16))
    assert d.max() <= 1 and (d == 0).mean() > 0.99

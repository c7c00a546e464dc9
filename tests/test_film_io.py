"""Film output (.hdr RGBE, .pfm) and EvaluateRMSE, on the CPU."""
import subprocess
import sys

import numpy as np
import pytest

from lightmetrica_amd import film_io


def test_rgbe_known_encodings():
    """Ward's shared-exponent encoding: frexp of the largest channel."""
    enc = film_io.rgbe_encode(np.array([[1.0, 1.0, 1.0], [0.0, 0.0, 0.0], [0.5, 0.25, 0.125], [1e-40, 0, 0]]))
    np.testing.assert_array_equal(enc[0], [128, 128, 128, 129])  # 1 = 0.5 * 2^1
    np.testing.assert_array_equal(enc[1], [0, 0, 0, 0])
    np.testing.assert_array_equal(enc[2], [128, 64, 32, 128])    # 0.5 = 0.5 * 2^0
    np.testing.assert_array_equal(enc[3], [0, 0, 0, 0])          # below 1e-32
    dec = film_io.rgbe_decode(enc)
    np.testing.assert_allclose(dec[0], [1.0, 1.0, 1.0], rtol=1 / 256)
    assert (dec[1] == 0).all()


def test_hdr_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    film = rng.lognormal(0.0, 2.0, (13, 17, 3)).astype(np.float32)
    film[0, 0] = 0.0
    p = film_io.write_hdr(tmp_path / "out", film)
    assert p.suffix == ".hdr"
    back = film_io.read_hdr(p)
    assert back.shape == film.shape
    # each channel within one part in 128 of the pixel's largest channel
    err = np.abs(back - film) / np.maximum(film.max(axis=2, keepdims=True), 1e-30)
    assert err.max() < 1.0 / 128
    assert (back[0, 0] == 0).all()
    # the header says top row first; row 0 of the film (bottom) is the last scanline
    raw = p.read_bytes()
    assert raw.startswith(b"#?RADIANCE\n") and b"-Y 13 +X 17\n" in raw
    last = np.frombuffer(raw[-17 * 4:], dtype=np.uint8).reshape(17, 4)
    np.testing.assert_array_equal(last, film_io.rgbe_encode(film[0]))


def test_hdr_reads_run_length_scanlines(tmp_path):
    """A hand-made new-style RLE file (2 2 hi lo, per-channel runs and dumps)."""
    w, h = 10, 2
    header = f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y {h} +X {w}\n".encode()
    rows = []
    pixels = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        pixels[y, :, 0] = 100 + y
        pixels[y, :, 1] = np.arange(w, dtype=np.uint8)
        pixels[y, :, 2] = 7
        pixels[y, :, 3] = 129
        line = bytes([2, 2, 0, w])
        line += bytes([128 + w, 100 + y])             # run
        line += bytes([w]) + bytes(range(w))          # dump
        line += bytes([128 + 4, 7, 128 + 6, 7])       # two runs
        line += bytes([128 + w, 129])
        rows.append(line)
    p = tmp_path / "rle.hdr"
    p.write_bytes(header + b"".join(rows))
    got = film_io.read_hdr(p)
    want = film_io.rgbe_decode(pixels)[::-1]
    np.testing.assert_array_equal(got, want)


def test_pfm_round_trip_is_exact(tmp_path):
    film = np.random.default_rng(2).standard_normal((5, 7, 3)).astype(np.float32)
    p = film_io.write_pfm(tmp_path / "a.pfm", film)
    np.testing.assert_array_equal(film_io.read_pfm(p), film)
    assert p.read_bytes().startswith(b"PF\n7 5\n-1.0\n")


def test_rmse_matches_the_reference_definition():
    a = np.arange(12, dtype=np.float64).reshape(2, 2, 3)
    assert film_io.rmse(a, a) == 0.0
    assert film_io.rmse(a + 0.25, a) == pytest.approx(0.25)
    # weight scales the second image: a - w * b
    assert film_io.rmse(a, a / 2, weight=2.0) == 0.0
    d = np.random.default_rng(3).standard_normal(300)
    assert film_io.rmse(d, np.zeros(300)) == pytest.approx(np.sqrt(np.mean(d ** 2)))
    with pytest.warns(UserWarning):
        assert film_io.rmse(a, a[:1]) == 0.0


def test_rmse_command_line(tmp_path):
    a = np.ones((4, 4, 3), np.float32)
    pa = film_io.write_pfm(tmp_path / "a.pfm", a)
    pb = film_io.write_pfm(tmp_path / "b.pfm", a * 1.5)
    out = subprocess.run([sys.executable, "-m", "lightmetrica_amd.film_io", "rmse", str(pa), str(pb)],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    assert float(out) == pytest.approx(0.5)

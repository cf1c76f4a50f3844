"""Pins the oracle's restatement of the output drivers (image/hdr/hdr.go:24-48 convertRGBToRGBE,
image/hdr/writer.go:62-107, builtin/driver/outputfloat.go:30-42).

Known answers are worked by hand from the Go source's float32 arithmetic on amd64: d = max(r, g, b)
by "greater than" tests; Frexp(d) = (n, e) with n in [0.5, 1); df = n * 255.999 / d; each byte is
the truncated int32 of x * df, low byte (NaN -> 0x80000000 -> 0); the exponent byte is e + 128.
"""
import math

import numpy as np
import pytest

import oracle

NAN, INF = float("nan"), float("inf")

KATS = [
    # (r, g, b) -> bytes, with the working
    ((1.0, 1.0, 1.0), (127, 127, 127, 129)),      # Frexp(1) = (0.5, 1): df = 127.9995
    ((0.5, 0.25, 0.0), (127, 63, 0, 128)),        # (0.5, 0): df = 255.999; 0.25 * df = 63.99975
    ((3.0, 1.5, 0.75), (191, 95, 47, 130)),       # (0.75, 2): df = 63.99975 -> 191.99925, 95.99963, 47.99981
    ((9e-7, 0.0, 0.0), (0, 0, 0, 0)),             # below the 1e-6 threshold
    ((-1.0, -2.0, -0.5), (0, 0, 0, 0)),           # max is negative: below the threshold
    ((NAN, 1.0, 0.0), (0, 0, 0, 128)),            # d = r = NaN ("g > NaN" is false): Frexp(NaN) = (NaN, 0)
    ((1.0, NAN, 0.5), (127, 0, 63, 129)),         # a NaN channel is never the max; its byte is 0
    ((2.0, -1.0, 0.5), (127, 193, 31, 130)),      # -63.99975 truncates to -63: low byte 0xC1
    ((INF, 1.0, 1.0), (0, 0, 0, 128)),            # Frexp(Inf) = (Inf, 0): df = Inf / Inf = NaN
    ((1e30, 0.0, 0.0), (201, 0, 0, 228)),         # 1e30 = 0.789 * 2^100: 0.789 * 255.999 = 201.96
]


@pytest.mark.parametrize("rgb,want", KATS)
def test_rgbe_known_answers(rgb, want):
    film = np.array(rgb, np.float32).reshape(1, 1, 3)
    assert tuple(oracle.film_to_rgbe(film)[0, 0]) == want


def test_rgbe_round_trip_bound():
    """Decoding as the reference's convertComponent does, byte / 256 * 2^(e - 128), gives every channel
    to within 2^e / 256 <= max / 128 below its value (the max channel truncates, never rounds up)."""
    g = np.random.default_rng(4)
    n = 20000
    film = (np.exp(g.uniform(-12, 12, (n, 1, 3))) * (g.random((n, 1, 3)) > 0.1)).astype(np.float32)
    enc = oracle.film_to_rgbe(film[::-1]).astype(np.int64)      # one-column film: rows reversed in the file
    e = enc[:, 0, 3]
    ok = e > 0
    dec = enc[:, 0, :3] / 256.0 * np.exp2((e - 128).astype(np.float64))[:, None]
    x = film[:, 0].astype(np.float64)
    d = x.max(1)
    assert np.all(d[~ok] < 1e-6)
    err = x[ok] - dec[ok]
    assert np.all(err >= -1e-12 * d[ok, None])
    assert np.all(err <= d[ok, None] / 128.0)


def test_hdr_file_layout(tmp_path):
    """Header text, the "+Y h +X w" resolution line, then w * h * 4 bytes with the bottom film row
    first and no run-length encoding."""
    h, w = 3, 5
    film = np.zeros((h, w, 3), np.float32)
    film[h - 1, 0] = (1.0, 1.0, 1.0)      # bottom-left pixel: the file's first pixel
    film[0, w - 1] = (3.0, 1.5, 0.75)     # top-right pixel: the file's last pixel
    p = tmp_path / "a.hdr"
    oracle.write_hdr(film, p)
    data = p.read_bytes()
    head = (b"#?RADIANCE\n# Created by Vermeer Light Tools (http://www.vermeerlt.com)\n"
            b"FORMAT=32-bit_rle_rgbe\n\n+Y 3 +X 5\n")
    assert data[: len(head)] == head and len(data) == len(head) + w * h * 4
    body = np.frombuffer(data[len(head):], np.uint8).reshape(h, w, 4)
    assert tuple(body[0, 0]) == (127, 127, 127, 129)
    assert tuple(body[h - 1, w - 1]) == (191, 95, 47, 130)
    assert body.sum() == 127 * 3 + 129 + 191 + 95 + 47 + 130


def test_float_file_is_little_endian_film_order(tmp_path):
    film = np.arange(2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3) * np.float32(0.37)
    film[1, 2, 1] = np.nan
    p = tmp_path / "a.float"
    oracle.write_float(film, p)
    assert p.read_bytes() == film.astype("<f4").tobytes()

"""Frame outputs of src/main.rs: the PNG dump (693-715) and the timing.txt line (1071-1073)."""
import struct
import zlib

import numpy as np

import temporalsetinversion_amd as T


def read_png(path):
    """Minimal PNG reader for 8-bit RGBA, filter 0 (what the library writes); checks every CRC."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    for h, w in ((1, 1), (3, 5), (200, 130), (257, 300)):   # > 64 KiB of scanlines: several blocks
        img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        p = tmp_path / f"{h}x{w}.png"
        T.write_png(p, img)
        assert np.array_equal(read_png(p), img)


def test_timing_line_matches_the_rust_format():
    # write!(file, "framenumber: {:?}, rootfinding_time_accumulated: {:?}, actual_frame_time: {:?}\n", ...)
    assert T.timing_line(0, 1.5, 16.667) == "framenumber: 0, rootfinding_time_accumulated: 1.5, actual_frame_time: 16.667\n"
    assert T.timing_line(12, 3.0, 0.0) == "framenumber: 12, rootfinding_time_accumulated: 3.0, actual_frame_time: 0.0\n"
    assert T._rust_f64_debug(1e-7) == "1e-7" and T._rust_f64_debug(1e16) == "1e16"
    assert T._rust_f64_debug(0.1 + 0.2) == "0.30000000000000004"

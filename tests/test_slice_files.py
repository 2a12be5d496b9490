"""Slice files in the reference's on-disk format (src/util.c:179-343):
frames of three little-endian uint64 {compressed size, chunk size, offset}
followed by a zstd frame (level 6) of at most 50 KiB of the slice's bytes,
in any order.  Host-only entry points: no device needed."""
import ctypes
import random
import struct

import numpy as np
import pytest

import tb8gen_amd as tbg

CHUNK = 50 * 1024
ZSTD_MAGIC = b"\x28\xb5\x2f\xfd"


def frames(path):
    out = []
    with open(path, "rb") as f:
        blob = f.read()
    pos = 0
    while pos < len(blob):
        csize, chunk, off = struct.unpack_from("<QQQ", blob, pos)
        out.append((csize, chunk, off, blob[pos + 24:pos + 24 + csize]))
        pos += 24 + csize
    assert pos == len(blob)
    return out


def bitmap_like(n, rng):
    # sparse bitmap words with runs, like a slice plane
    a = np.zeros(n, dtype=np.uint8)
    idx = rng.integers(0, max(n, 1), size=max(n // 40, 1))
    a[idx] = rng.integers(1, 256, size=idx.size, dtype=np.uint8)
    return a


@pytest.mark.parametrize("size", [1, 64, CHUNK - 8, CHUNK, CHUNK + 8, 7 * CHUNK + 3, 1 << 20])
def test_round_trip_and_frame_layout(tmp_path, size):
    rng = np.random.default_rng(size)
    data = bitmap_like(size, rng)
    p = tmp_path / "slice"
    tbg.file_write(p, data)
    fr = frames(p)
    assert len(fr) == (size + CHUNK - 1) // CHUNK
    covered = 0
    for csize, chunk, off, payload in fr:
        assert chunk <= CHUNK and off % CHUNK == 0 and off + chunk <= size
        assert payload[:4] == ZSTD_MAGIC and len(payload) == csize
        covered += chunk
    assert covered == size
    assert np.array_equal(tbg.file_read(p, size), data)


def test_frames_in_any_order(tmp_path):
    """The reference's writer threads emit frames out of order; the reader
    places each by its offset."""
    size = 9 * CHUNK + 100
    data = bitmap_like(size, np.random.default_rng(1))
    p = tmp_path / "a"
    tbg.file_write(p, data)
    fr = frames(p)
    random.Random(3).shuffle(fr)
    q = tmp_path / "b"
    with open(q, "wb") as f:
        for csize, chunk, off, payload in fr:
            f.write(struct.pack("<QQQ", csize, chunk, off) + payload)
    assert np.array_equal(tbg.file_read(q, size), data)


def test_damaged_files_are_errors(tmp_path):
    size = 3 * CHUNK
    data = bitmap_like(size, np.random.default_rng(2))
    p = tmp_path / "s"
    tbg.file_write(p, data)
    blob = p.read_bytes()
    L = tbg.lib()
    buf = np.zeros(size, dtype=np.uint8)
    ptr = buf.ctypes.data_as(ctypes.c_void_p)
    # truncated
    (tmp_path / "t").write_bytes(blob[:-10])
    assert L.tbg_file_read(str(tmp_path / "t").encode(), ptr, size) == -8
    # a chunk size beyond 50 KiB
    bad = bytearray(blob)
    struct.pack_into("<Q", bad, 8, CHUNK + 1)
    (tmp_path / "h").write_bytes(bytes(bad))
    assert L.tbg_file_read(str(tmp_path / "h").encode(), ptr, size) == -8
    # corrupt compressed bytes
    bad = bytearray(blob)
    bad[24 + 8] ^= 0xFF
    (tmp_path / "c").write_bytes(bytes(bad))
    assert L.tbg_file_read(str(tmp_path / "c").encode(), ptr, size) == -8
    # more bytes expected than the file holds, and a missing file
    assert L.tbg_file_read(str(p).encode(), ptr, size - 1) == -8
    assert L.tbg_file_read(str(tmp_path / "none").encode(), ptr, size) == -8
    assert L.tbg_strerror(-8)


def test_duplicated_and_misplaced_frames_are_errors(tmp_path):
    """Frames must cover disjoint whole chunks: a file with the first frame
    twice (and the last chunk missing), or with a frame moved off its chunk
    boundary, is rejected instead of read with part of the slice unwritten
    or decompressed twice into the same bytes."""
    size = 3 * CHUNK
    data = bitmap_like(size, np.random.default_rng(4))
    p = tmp_path / "s"
    tbg.file_write(p, data)
    fr = frames(p)
    L = tbg.lib()
    buf = np.zeros(size, dtype=np.uint8)
    ptr = buf.ctypes.data_as(ctypes.c_void_p)

    def write(name, frs):
        q = tmp_path / name
        with open(q, "wb") as f:
            for csize, chunk, off, payload in frs:
                f.write(struct.pack("<QQQ", csize, chunk, off) + payload)
        return str(q).encode()

    # every header alone is in bounds; the byte counts add up to the size
    assert L.tbg_file_read(write("dup", [fr[0], fr[1], fr[0]]), ptr, size) == -8
    assert L.tbg_file_read(write("ok", [fr[2], fr[0], fr[1]]), ptr, size) == 0
    assert np.array_equal(buf, data)
    csize, chunk, off, payload = fr[1]
    assert L.tbg_file_read(write("mis", [fr[0], (csize, chunk, off + 8, payload), fr[2]]),
                           ptr, size) == -8
    # a short chunk inside the slice (only the last may be short)
    assert L.tbg_file_read(write("short", [fr[0], (csize, chunk - 8, off, payload), fr[2]]),
                           ptr, size) == -8

"""Image writers (ahip_write_png / ahip_write_exr, anacapa_amd/csrc/film_io.cpp).

PNG is read back by an independent decoder (Pillow).  No OpenEXR reader exists in this image, so
the EXR file is read back by a reader written here from the OpenEXR file-layout specification
(magic, version, typed header attributes, offset table, uncompressed scanline blocks); agreement
with OpenImageIO's ZIP-compressed output of the reference is not checked (parity unpinned there).
"""
import os
import struct

import numpy as np
import pytest

from anacapa_amd import _abi as A
from anacapa_amd.integrator import BackendError, write_exr, write_png


def test_png_round_trip(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, size=(37, 53, 3), dtype=np.uint8)
    path = tmp_path / "frame.png"
    write_png(path, img)
    assert not os.path.exists(str(path) + ".writing.png")        # written aside, renamed into place
    with Image.open(path) as im:
        assert im.mode == "RGB" and im.size == (53, 37)
        assert np.array_equal(np.asarray(im), img)


def _read_exr(path):
    b = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", b, 0)
    pos, attrs = 8, {}
    while b[pos] != 0:
        end = b.index(b"\0", pos)
        name = b[pos:end].decode()
        tend = b.index(b"\0", end + 1)
        typ = b[end + 1:tend].decode()
        (size,) = struct.unpack_from("<i", b, tend + 1)
        attrs[name] = (typ, b[tend + 5:tend + 5 + size])
        pos = tend + 5 + size
    pos += 1
    chans, cb = [], attrs["channels"][1]
    p = 0
    while cb[p] != 0:
        end = cb.index(b"\0", p)
        ptype, = struct.unpack_from("<i", cb, end + 1)
        chans.append((cb[p:end].decode(), ptype))
        p = end + 1 + 16
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", b, pos)
    data = {c: np.zeros((h, w), np.float32) for c, _ in chans}
    for off in offsets:
        y, size = struct.unpack_from("<ii", b, off)
        assert size == 4 * w * len(chans)
        row = np.frombuffer(b, "<f4", w * len(chans), off + 8).reshape(len(chans), w)
        for k, (c, _) in enumerate(chans):
            data[c][y - y0] = row[k]
    return magic, version, attrs, chans, data


def test_exr_layout_and_values(tmp_path):
    rng = np.random.default_rng(2)
    h, w = 9, 14
    beauty, albedo, normal = (rng.normal(size=(h, w, 3)).astype(np.float32) for _ in range(3))
    path = tmp_path / "frame.exr"
    write_exr(path, {"": beauty, "albedo.": albedo, "normals.": normal})
    magic, version, attrs, chans, data = _read_exr(path)
    assert magic == 20000630 and version == 2
    names = [c for c, _ in chans]
    assert names == sorted(names) == ["B", "G", "R", "albedo.B", "albedo.G", "albedo.R", "normals.B",
                                      "normals.G", "normals.R"]
    assert all(t == 2 for _, t in chans)                             # FLOAT
    assert attrs["compression"] == ("compression", b"\0")
    assert attrs["dataWindow"] == attrs["displayWindow"] == ("box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    assert attrs["lineOrder"][1] == b"\0"
    for prefix, img in (("", beauty), ("albedo.", albedo), ("normals.", normal)):
        for k, c in enumerate("RGB"):
            assert np.array_equal(data[prefix + c], img[..., k]), prefix + c


def test_writer_errors(tmp_path):
    import ctypes as C
    lib = A.load_library()
    img = np.zeros((4, 4, 3), np.float32)
    pre = (C.c_char_p * 2)(b"x.", b"x.")                            # the same channel names twice
    ptrs = (C.c_void_p * 2)(img.ctypes.data, img.ctypes.data)
    assert lib.ahip_write_exr(str(tmp_path / "dup.exr").encode(), 4, 4, 2, pre, ptrs) == A.AHIP_E_INVALID_ARG
    assert not os.path.exists(tmp_path / "dup.exr")
    with pytest.raises(BackendError):
        write_png(tmp_path / "missing" / "x.png", np.zeros((2, 2, 3), np.uint8))

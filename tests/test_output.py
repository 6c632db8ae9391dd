"""Output files (CPU): the TIFF writer reproduces WriteTifMonochrome16's layout
(general.cpp:1313-1470) and the ImageDescription of gazoDoc.cpp:1986-1988."""
import os
import struct

import numpy as np
import pytest

from recview_amd import _lib, output

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")


def test_image_description_format():
    s = output.image_description(1800, 0.5, 12.25, -3.5, 1024.37, 0)
    assert s == "1800\t0.500000\t12.250000\t-3.500000\t1024.370000\t1\t55.111697\t0.461227 "


@pytest.mark.parametrize("desc", ["abc", "1800\t0.500000 ", "x" * 300, None])
def test_tif_layout(tmp_path, desc):
    H, W = 7, 5
    img = np.arange(H * W, dtype=np.int32).reshape(H, W) * 1000 - 3000
    img[0, 0] = 70000                              # wraps like the reference's (unsigned short) cast
    p = tmp_path / "s.tif"
    output.write_tif16(p, img, desc)
    data = p.read_bytes()
    d = {"abc": "abc    ", None: "tif image"}.get(desc, desc if desc is None or len(desc) < 200 else desc[:199])
    nd = len(d) + 1
    na = len("SP-uCT, gazoView") + 1
    expected = 8 + 2 * H * W + 2 + 12 * 12 + 4 + nd + (nd & 1) + 8 * H + na + (na & 1)
    assert len(data) == expected
    assert data[:4] == b"II*\x00"
    (ifd,) = struct.unpack_from("<I", data, 4)
    assert ifd == 8 + 2 * H * W
    tags = [struct.unpack_from("<HHII", data, ifd + 2 + 12 * i) for i in range(12)]
    assert [t[0] for t in tags] == [0x100, 0x101, 0x102, 0x103, 0x106, 0x10E, 0x111, 0x115, 0x116, 0x117,
                                    0x11C, 0x13B]
    pix = np.frombuffer(data[8:8 + 2 * H * W], "<u2").reshape(H, W)
    ref = np.where(img > 0, img, 0).astype(np.int64) & 0xFFFF
    assert np.array_equal(pix, ref)
    back, dd, art = output.read_tif16(p)
    assert np.array_equal(back, ref) and dd == d and art == "SP-uCT, gazoView"
    assert struct.unpack_from("<I", data, ifd + 2 + 12 * 12)[0] == 0          # single directory

"""HDF5 scan metadata (APS DXchange): the minimal reader against files written by
the HDF5 library (tests/golden/make_hdf5_fixtures.py), and the scan log and
exclusion list RecView derives from them (gazoDoc.cpp:2229-2312, 2438-2468)."""
import json
import os

import numpy as np
import pytest

from recview_amd import logfile as LF
from recview_amd.hdf5 import H5Error, H5File

GOLD = os.path.join(os.path.dirname(__file__), "golden")
META = json.load(open(os.path.join(GOLD, "hdf5_fixtures.json")))


@pytest.mark.parametrize("name", sorted(META))
def test_reader_matches_library_written_files(name):
    m = META[name]
    with H5File(os.path.join(GOLD, name)) as f:
        assert f.shape("/exchange/data")[0] == m["frames"]
        assert f.exists("/exchange") and not f.exists("/exchange/nothing")
        if m["theta"] is None:
            assert not f.exists("/exchange/theta")
        else:
            t = f.read("/exchange/theta")
            assert np.array_equal(t.astype(np.float64), np.array(m["theta"]))


def test_reader_group_with_several_datasets_and_bytes_input():
    raw = open(os.path.join(GOLD, "dxchange_step_f32.h5"), "rb").read()
    f = H5File(raw)
    assert f.shape("/exchange/data_dark") == (2, 4, 6)
    assert f.shape("/exchange/data_white") == (3, 4, 6)
    with pytest.raises(H5Error):
        f.read("/exchange/data")              # chunked: dimensions only
    with pytest.raises(KeyError):
        f.shape("/exchange/theta2")
    with pytest.raises(H5Error):
        H5File(b"not an hdf5 file" * 8)


def test_step_scan_log():
    for name in ("dxchange_step_f32.h5", "dxchange_step_f64_latest.h5"):
        theta = np.array(META[name]["theta"])
        log = LF.load_hdf5_log(os.path.join(GOLD, name))
        n = len(theta)
        assert log.step_scan and log.hdf5_frames == META[name]["frames"]
        assert log.len_sinogr == n + 3                          # white, n samples, white, dark
        assert np.array_equal(log.fdeg[1:n + 1], theta.astype(np.float32))
        assert log.fdeg[0] == log.fdeg[1] - 1 and log.fdeg[n + 1] == log.fdeg[n] + 1
        assert list(log.binc) == [LF.BINC_WHITE] + [LF.BINC_SAMPLE] * n + [LF.BINC_WHITE]
        assert log.fname[1] == "00001" and np.array_equal(log.fexp, log.fdeg)
        excl, first, last = LF.frames_to_exclude(log)
        assert excl.split() == [] and first == 0 and last == n - 1      # less than 180 degrees


def test_fly_scan_log_and_exclusions():
    log = LF.load_hdf5_log(os.path.join(GOLD, "dxchange_fly.h5"))
    N = META["dxchange_fly.h5"]["frames"]
    assert not log.step_scan and log.len_sinogr == N + 3
    d = np.diff(log.fdeg[1:N + 1].astype(np.float64))
    assert np.allclose(d, 180.0 / (N - 1), rtol=0, atol=4e-5)      # evenly spread (float32 angles)
    assert log.fdeg[N - 1] == np.float32(180.0)                     # entry i: 180*i/(N-1) degrees
    # the frame at 180 degrees repeats the first view and is excluded; the
    # averaged white frame b0 is listed for fly scans
    excl, first, last = LF.frames_to_exclude(log)
    assert excl.split() == ["s19", "b00"] and first == 0 and last == N - 2


def test_step_scan_exclusions_with_leading_frames_and_overturn():
    """Frames in /exchange/data before the theta list are excluded by number, and
    sample angles beyond 179.99 degrees from the first are excluded with the
    data-block offset."""
    theta = np.array([10.0, 70.0, 130.0, 189.99, 200.0], np.float32)
    n = len(theta) + 2
    fdeg = np.zeros(n, np.float32)
    fdeg[1:n - 1] = theta
    fdeg[0], fdeg[n - 1] = fdeg[1] - 1, fdeg[n - 2] + 1
    binc = np.full(n, LF.BINC_SAMPLE, np.int8)
    binc[0] = binc[n - 1] = LF.BINC_WHITE
    log = LF.ScanLog(["%05d" % i for i in range(n)], fdeg.copy(), fdeg, binc, n + 1, 8, True)
    excl, first, last = LF.frames_to_exclude(log)
    # 3 leading frames (8 in the data block, 5 angles); 189.99 and 200 are >= 179.99 past 10
    assert excl.split() == ["s0", "s1", "s2", "s6", "s7"]
    assert first == 3 and last == 5


def test_reader_rejects_cyclic_continuation_blocks():
    """An object header whose continuation message points back at its own block
    ends in an error instead of an endless walk."""
    import struct
    raw = open(os.path.join(GOLD, "dxchange_step_f32.h5"), "rb").read()
    f = H5File(raw)
    so, sl = f.so, f.sl
    a = len(raw)
    body = (a + 16 - f.base).to_bytes(so, "little") + (8 + so + sl).to_bytes(sl, "little")
    msg = struct.pack("<HHB3x", 0x10, len(body), 0) + body
    head = struct.pack("<BBHII4x", 1, 0, 5, 1, len(msg))
    g = H5File(raw + head + msg)
    with pytest.raises(H5Error, match="cycle"):
        g._messages(a - g.base)

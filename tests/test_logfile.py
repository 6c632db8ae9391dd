"""Scan-log ingestion and frame selection on the reference's own angle
schedules (docs/output450.log, docs/output1800offset.log, copied as fixtures)."""
import os

import numpy as np
import pytest

from recview_amd import logfile as LF

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _used(binc):
    return np.nonzero(((binc & LF.BINC_SAMPLE) != 0) & ((binc & LF.BINC_NOUSE) == 0))[0]


def test_output450_log():
    log = LF.load_output_log(os.path.join(GOLD, "output450.log"))
    assert len(log.fname) == 454 and log.len_sinogr == 455          # blank lines skipped
    assert log.fname[0] == "00001" and log.fname[-1] == "00454"
    assert (log.binc == LF.BINC_SAMPLE).sum() == 451 and (log.binc == LF.BINC_WHITE).sum() == 3
    assert log.fdeg.dtype == np.float32 and log.fdeg[0] == 0 and log.fdeg[1] == np.float32(0.4)
    excl, first, last = LF.frames_to_exclude(log)
    # 0.4-degree steps from 0.4 to 180.4: the sample after the closing white
    # frames (180.4) is excluded, and so is the one at 180.0, because the
    # reference compares it in double with fdegEnd - notch = 180.4f - 0.4f
    # = 179.9999939 (gazoDoc.cpp:2492)
    assert excl.split() == ["s453", "s450"]
    assert (first, last) == (1, 449)
    used = _used(LF.apply_exclusions(log, excl))
    assert len(used) == 449 and used[0] == 1 and used[-1] == 449
    assert np.isclose(log.fdeg[used].max() - log.fdeg[used].min(), 179.2)


def test_output1800offset_log_360():
    log = LF.load_output_log(os.path.join(GOLD, "output1800offset.log"))
    excl, first, last = LF.frames_to_exclude(log)
    assert excl.split() == ["s3603"]                                 # the sample after the white frames
    used = _used(LF.apply_exclusions(log, excl))
    span = float(log.fdeg[used].max() - log.fdeg[used].min())
    assert np.isclose(span, 359.9) and len(used) == 3600              # one full turn (360 deg scan)
    assert first == used[0] == 1 and last == used[-1] == 3600


def test_irregular_and_empty_logs():
    with pytest.raises(LF.LogError) as e:
        LF.load_output_log("00001 0.1 0.0 0\n00002 abc 0.1 1\n")
    assert e.value.code == 21120403
    with pytest.raises(LF.LogError) as e:
        LF.load_output_log("\n\n")
    assert e.value.code == 21120401


def test_pulse_angles_and_negative_start():
    log = LF.load_output_log("00001 0.1 -5.0 0\n00002 0.2 45000 1\n00003 0.3 90000 1\n")
    assert log.fdeg[0] == np.float32(np.float64(np.float32(-0.1)) / 500.0) and log.fdeg[2] == 180


def test_bare_frame_number_replaced_by_neighbours():
    log = LF.load_output_log("".join(f"{i:05d} {i * 0.1:.5f} {i * 1.0:.4f} {0 if i in (1, 9) else 1}\n"
                                     for i in range(1, 10)))
    rows = np.arange(9 * 2 * 3, dtype=np.int16).reshape(9, 2, 3) * 7
    orig = rows.copy()
    binc = LF.apply_exclusions(log, " 3 s5 0", rows)
    # frame 3 (sample) = trunc((frame 2 + frame 4) / 2); frame 5 -> NOUSE|SKIP; frame 0 (white) -> NOUSE
    assert np.array_equal(rows[3], ((orig[2].astype(int) + orig[4]) // 2).astype(np.int16))
    assert binc[5] == LF.BINC_SAMPLE | LF.BINC_NOUSE | LF.BINC_SKIP
    assert binc[0] == LF.BINC_NOUSE and binc[3] == LF.BINC_SAMPLE

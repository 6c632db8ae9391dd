"""Scan logs and frame selection (the input side of a reconstruction).

load_output_log   -- SPring-8 output.log (gazoDoc.cpp:2298-2380): one line per
                     frame, "name exposure angle flag"; flag '1' = sample frame,
                     anything else = incident (white) frame; the dark frame is
                     appended by the caller (iLenSinogr = frames + 1).
load_hdf5_log     -- APS DXchange HDF5 (gazoDoc.cpp:2229-2312): a step scan
                     takes its angles from /exchange/theta, a fly scan (no
                     usable theta) spreads 180 degrees over the frames of
                     /exchange/data; both get a white entry before and after.
frames_to_exclude -- the default exclusion list for a log (gazoDoc.cpp:2411-2520):
                     sample frames beyond one 180 (or 360) degree turn, as
                     tokens " sNNNN" (HDF5: frame numbers of the data block).
apply_exclusions  -- turns an exclusion list into bInc flags and, for bare frame
                     numbers of sample frames, replaces the frame by the mean of
                     its nearest kept neighbours (gazoDoc.cpp:3372-3467).
"""
import math
from dataclasses import dataclass, field

import numpy as np

BINC_WHITE, BINC_SAMPLE, BINC_NOUSE, BINC_SKIP = 0, 1, 2, 4     # StdAfx.h:28-31


class LogError(ValueError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} ({code})")
        self.code = code


@dataclass
class ScanLog:
    fname: list
    fexp: np.ndarray          # float32 exposure times
    fdeg: np.ndarray          # float32 angles (degrees)
    binc: np.ndarray          # int8 flags
    len_sinogr: int = field(default=0)   # frames + dark
    hdf5_frames: int = field(default=0)  # HDF5 logs: frames in /exchange/data
    step_scan: bool = field(default=False)   # HDF5 logs: angles from /exchange/theta


def _scan_line(line):
    """sscanf("%s %f %f %c"): returns (name, exp, deg, flag char) or None."""
    parts = line.split()
    if len(parts) < 4:
        return None
    try:
        return parts[0][:19], np.float32(float(parts[1])), np.float32(float(parts[2])), parts[3][0]
    except ValueError:
        return None


def load_output_log(source):
    """Parses an output.log (path or text)."""
    text = source
    if "\n" not in source and not source.lstrip().startswith("0"):
        with open(source, "rt") as f:
            text = f.read()
    lines = [ln for ln in text.splitlines() if ln != ""]
    if not lines:
        raise LogError(21120401, "empty output.log")
    fname, fexp, fdeg, binc = [], [], [], []
    for ln in lines:
        r = _scan_line(ln)
        if r is None:
            raise LogError(21120403, f"irregular line in output.log: {ln!r}")
        fname.append(r[0])
        fexp.append(r[1])
        fdeg.append(r[2])
        binc.append(BINC_SAMPLE if r[3] == "1" else BINC_WHITE)
    fexp = np.array(fexp, np.float32)
    fdeg = np.array(fdeg, np.float32)
    if fdeg[0] < -0.1:
        fdeg[0] = np.float32(-0.1)
    if abs(fdeg[-1]) > 10000:            # angles given in pulses
        fdeg = (fdeg.astype(np.float64) / 500.0).astype(np.float32)   # float /= double
    return ScanLog(fname, fexp, fdeg, np.array(binc, np.int8), len(lines) + 1)


def load_hdf5_log(path):
    """Scan log of an APS DXchange HDF5 file (LoadLogFile, gazoDoc.cpp:2229-2312)."""
    from .hdf5 import H5File, H5Error
    with H5File(path) as f:
        try:
            dims = f.shape("/exchange/data")
        except KeyError:
            raise LogError(16052221, "no /exchange/data in the HDF5 file") from None
        nframes = int(dims[0])
        theta = None
        try:
            if f.exists("/exchange/theta") and len(f.shape("/exchange/theta")) == 1:
                theta = f.read("/exchange/theta")
        except H5Error:
            theta = None                  # unreadable theta: handled as a fly scan, as the reference
    if theta is not None:
        n = len(theta) + 2
        fdeg = np.zeros(n, np.float32)
        fdeg[1:n - 1] = theta.astype(np.float32)              # (float) of float / double elements
    else:
        n = nframes + 2
        fdeg = np.zeros(n, np.float32)
        i = np.arange(1, n - 1)
        fdeg[1:n - 1] = (np.float32(180) * i.astype(np.float32)) / np.float32(n - 3)   # 180.f*i/(ipos-2-1)
    fdeg[0] = fdeg[1] - np.float32(1)
    fdeg[n - 1] = fdeg[n - 2] + np.float32(1)
    binc = np.full(n, BINC_SAMPLE, np.int8)
    binc[0] = binc[n - 1] = BINC_WHITE
    names = ["%05d" % i for i in range(n)]
    return ScanLog(names, fdeg.copy(), fdeg, binc, n + 1, nframes, theta is not None)


def frames_to_exclude(log):
    """(exclusion string, first sample frame, last sample frame) for a log."""
    if log.hdf5_frames or log.step_scan:
        return _frames_to_exclude_hdf5(log)
    fdeg, binc = log.fdeg, log.binc
    ipos = len(fdeg)
    idigit = int(math.log10(ipos)) + 1
    fmt = " s%0" + str(idigit) + "d"
    sample = (binc & BINC_SAMPLE) != 0
    idx = np.nonzero(sample)[0]
    deg_start = float(fdeg[idx[0]]) if len(idx) else 0.0
    deg_end = float(fdeg[idx[-1]]) if len(idx) else 180.0
    span = np.float32(np.float32(deg_end) - np.float32(deg_start)) - np.float32(360)
    d_end = 359.99 if abs(float(span)) < 10 else 179.99
    notch = 0.1
    for i in range(1, ipos):
        if fdeg[i] != fdeg[i - 1]:
            notch = abs(float(np.float32(fdeg[i] - fdeg[i - 1])))    # float difference, then fabs
            break
    out = " "
    first, last = 0, 0
    for i in range(ipos - 1, 0, -1):
        if sample[i]:
            if float(fdeg[i]) <= deg_end - notch:
                last = i
                break
            out += fmt % i
    fl_end = float(fdeg[last]) if last else deg_end
    for i in range(ipos):
        if sample[i]:
            if fl_end - float(fdeg[i]) >= d_end:
                out += fmt % i
            else:
                first = i
                break
    return out, first, last


def _sample_span(log):
    ipos = log.len_sinogr - 1
    sample = (log.binc[:ipos] & BINC_SAMPLE) != 0
    idx = np.nonzero(sample)[0]
    start = log.fdeg[idx[0]] if len(idx) else np.float32(0)
    end = log.fdeg[idx[-1]] if len(idx) else np.float32(180)
    span = np.float32(np.float32(end) - np.float32(start)) - np.float32(360)
    return ipos, sample, np.float32(start), (359.99 if abs(float(span)) < 10 else 179.99)


def _frames_to_exclude_hdf5(log):
    """The HDF5 branch of SetFramesToExclude (gazoDoc.cpp:2438-2468): frame numbers
    of the data block; a fly scan also lists the averaged white frame (" b0")."""
    ipos, sample, start, d_end = _sample_span(log)
    fmt = " s%0" + str(int(math.log10(ipos)) + 1) + "d"
    out, first, last = " ", 0, 0
    if log.step_scan and log.hdf5_frames > 0:
        first = log.hdf5_frames - (ipos - 2)
        for i in range(first):
            out += fmt % i
        for i in range(1, ipos):
            if not sample[i]:
                continue
            if float(np.float32(log.fdeg[i] - start)) >= d_end:
                out += fmt % (i - 1 + first)
            else:
                last = i - 1 + first
    else:
        for i in range(ipos):
            if sample[i] and log.fdeg[i] == start:
                first = i - 1
                break
        for i in range(1, ipos):
            if sample[i]:
                if float(np.float32(log.fdeg[i] - start)) >= d_end:
                    out += fmt % (i - 1)
                else:
                    last = i - 1
        out += (" b%0" + str(int(math.log10(ipos)) + 1) + "d") % 0
    return out, first, last


def apply_exclusions(log, excl, rows=None):
    """Applies an exclusion list to a copy of log.binc.  rows (optional):
    int16 frames [frames][mult][W] of one sinogram, modified in place where a
    bare frame number asks for replacement by its kept neighbours."""
    binc = log.binc.copy()
    isino = log.len_sinogr
    tokens = excl.split()

    def listed(i):
        for t in tokens:
            if t[0] in "hbw":
                continue
            if int(t[1:] if t[0] == "s" else t) == i:
                return True
        return False

    for t in tokens:
        if t[0] in "hbw":
            continue
        if t[0] == "s":
            i = int(t[1:])
            if 0 <= i < isino - 1 and binc[i] & BINC_SAMPLE:
                binc[i] |= BINC_NOUSE | BINC_SKIP
            continue
        i = int(t)
        if not 0 <= i < isino - 1:
            continue
        if not binc[i] & BINC_SAMPLE:
            binc[i] |= BINC_NOUSE
            continue
        kept = lambda j: (binc[j] & BINC_SAMPLE) and not (binc[j] & BINC_NOUSE) and not listed(j)
        s0 = next((j for j in range(i - 1, -1, -1) if kept(j)), -1)
        s1 = next((j for j in range(i + 1, isino - 1) if kept(j)), -1)
        if s0 < 0 and s1 < 0:
            binc[i] |= BINC_NOUSE | BINC_SKIP
        elif rows is not None:
            acc = np.zeros(rows.shape[1:], np.int64)
            n = 0
            for j in (s0, s1):
                if j >= 0:
                    acc += rows[j].astype(np.int64)
                    n += 1
            rows[i] = (np.trunc(acc / n)).astype(np.int16)
    return binc

"""CudaSinogramHost (sinogram generation on the GPU) bit-exact against the C
restatement of GenerateSinogramThread (general.cpp:2377-2532)."""
import numpy as np
import pytest

from oracle import pyoracle as O
from recview_amd import sinogram

pytestmark = pytest.mark.gpu


def _scan(seed, nframes=40, mult=2, W=300, dark=120):
    rng = np.random.default_rng(seed)
    flat = rng.uniform(20000, 40000, (mult, W))
    binc = np.ones(nframes, np.int8)
    binc[[i for i in (0, 13, 27) if i < nframes - 1]] = 0     # incident frames
    fexp = np.cumsum(rng.uniform(0.5, 1.5, nframes)).astype(np.float32)
    fdeg = np.linspace(0, 180, nframes, endpoint=False).astype(np.float32)
    frames = np.empty((nframes, mult, W))
    for i in range(nframes - 1):
        drift = 1 + 0.05 * np.sin(fexp[i])                  # slowly varying beam
        if binc[i] == 0:
            frames[i] = flat * drift + dark
        else:
            p = rng.uniform(0, 4, (mult, W))
            frames[i] = rng.poisson(flat * drift * np.exp(-p)) + dark
    frames[-1] = dark + rng.integers(-5, 6, (mult, W))
    frames[5, 0, :7] = dark + 10                            # below the pixel floor
    return np.clip(frames, 0, 65535).astype(np.uint16), binc, fexp, fdeg


@pytest.mark.parametrize("case", ["plain", "zernike", "8bit", "skip_nouse", "drift_list", "drift_params",
                                  "step", "binning"])
def test_sinogram_bit_exact(case):
    frames, binc, fexp, fdeg = _scan(7 + len(case))
    kw = {}
    if case == "zernike":
        kw["flags"] = 2
    elif case == "8bit":
        kw["is8bit"] = True
    elif case == "skip_nouse":
        binc[4] |= 4
        binc[13] |= 2        # a NOUSE incident frame: skipped as i0, still found as i1
    elif case == "drift_list":
        kw["flags"] = 4
        kw["pidrift"] = np.random.default_rng(1).integers(-9, 10, len(binc))
    elif case == "drift_params":
        kw["flags"] = 8
        kw["drift"] = (8, 30, 6.4, -3.2, 1)
    elif case == "step":
        kw["start"], kw["step"] = 1, 3
    elif case == "binning":
        kw["flags"] = 12
        kw["drift"] = (8, 30, 9.0, 4.0, 0)
        kw["pidrift"] = np.random.default_rng(2).integers(-9, 10, len(binc))
    interp = 0 if case == "binning" else 2
    got, st = sinogram.generate(frames, binc, fexp, fdeg, iInterpolation=interp, **kw)
    binning = 4 if case == "binning" else 1
    ref, rst = O.sinogram(frames, binc, fexp, fdeg, binning=binning, **kw)
    assert rst == 0 and st == 0
    assert np.array_equal(got, ref)
    assert (got != frames.view(np.int16)).any()


def test_sinogram_error_without_incident_frame():
    frames, binc, fexp, fdeg = _scan(3, nframes=6)
    binc[:] = 1
    got, st = sinogram.generate(frames, binc, fexp, fdeg)
    assert st == 2      # RECONST_INFO_ERROR

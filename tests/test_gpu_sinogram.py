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


@pytest.mark.parametrize("case", ["plain", "drift_list", "skip_nouse"])
def test_device_frames_int16_equals_host(case):
    import torch
    frames, binc, fexp, fdeg = _scan(31 + len(case))
    kw = {}
    if case == "drift_list":
        kw = {"flags": 4, "pidrift": np.random.default_rng(5).integers(-9, 10, len(binc))}
    elif case == "skip_nouse":
        binc[4] |= 4
        binc[13] |= 2
    host, st = sinogram.generate(frames, binc, fexp, fdeg, **kw)
    d = torch.from_numpy(frames.view(np.int16)).cuda()
    dev, n = sinogram.generate_device(d, binc, fexp, fdeg, out_float=False, **kw)
    torch.cuda.synchronize()
    dev = dev.cpu().numpy()
    conv = np.zeros(frames.shape[0], bool)
    conv[:-1] = (binc[:-1] & 1) != 0
    assert n == conv.sum() * frames.shape[1]
    assert np.array_equal(dev[conv], host[conv])


def test_device_frames_float_rows():
    """Float rows are the unquantised line integrals of the int16 rows."""
    import torch
    frames, binc, fexp, fdeg = _scan(41)
    d = torch.from_numpy(frames.view(np.int16)).cuda()
    f, _ = sinogram.generate_device(d, binc, fexp, fdeg, out_float=True)
    i16, _ = sinogram.generate_device(d, binc, fexp, fdeg, out_float=False)
    torch.cuda.synchronize()
    f, i16 = f.cpu().numpy(), i16.cpu().numpy().astype(np.float64)
    conv = (binc[:-1] & 1) != 0
    ff = f[:-1][conv].astype(np.float64) * 3000
    q = i16[:-1][conv]
    # int16 = (short)(3000 p + 0.5): truncation toward zero, also for negative p;
    # the float rows round p to FP32 first, so a borderline value may land one off
    expect = np.trunc(ff + 0.5)
    assert np.abs(expect - q).max() <= 1
    assert np.mean(expect != q) < 1e-3
    assert (np.abs(f[:-1][~conv]) == 0).all()                 # incident / dark rows untouched

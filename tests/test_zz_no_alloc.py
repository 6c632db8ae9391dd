"""RecviewOutputStageNoAlloc (the allocation-free output stage) against
RecviewOutputStage through the C ABI: equal uint16 images and equal (min, max)
on two calls in a row, the second on a smaller image (the shared per-GPU
partials keep the first call's contents beyond the second's), int32 and
float32 images; sizes beyond the limit are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

from recview_amd import _lib, output

pytestmark = pytest.mark.gpu


def _no_alloc(img, center, interp):
    N = img.shape[0]
    out = torch.empty(N, N, dtype=torch.int16, device="cuda")
    rng = torch.empty(2, dtype=img.dtype, device="cuda")
    s = torch.cuda.current_stream()
    rc = _lib.lib().RecviewOutputStageNoAlloc(C.c_void_p(img.data_ptr()), int(img.dtype == torch.float32), N,
                                              float(center), int(interp), C.c_void_p(out.data_ptr()),
                                              C.c_void_p(rng.data_ptr()), C.c_void_p(s.cuda_stream))
    assert rc == 0, _lib.lib().RecviewLastMessage()
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint16), rng.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_no_alloc_equals_output_stage(dtype):
    rs = np.random.default_rng(77)
    for N, interp, center, lo, hi in ((1500, 2, 750.3, -40000, 90000), (300, 3, 75.2, 5, 9000)):
        print(f"case N={N} dtype={np.dtype(dtype).name}", flush=True)
        img = torch.from_numpy(rs.integers(lo, hi, (N, N)).astype(dtype)).cuda()
        ref, rmin, rmax = output.output_stage(img, center, interp)
        torch.cuda.synchronize()
        got, rng = _no_alloc(img, center, interp)
        assert np.array_equal(got, ref.cpu().numpy().view(np.uint16)), f"N={N}"
        assert (float(rng[0]), float(rng[1])) == (rmin, rmax), f"N={N}"


def test_no_alloc_refuses_bad_sizes():
    L = _lib.lib()
    img = torch.zeros(4, 4, dtype=torch.int32, device="cuda")
    out = torch.empty(4, 4, dtype=torch.int16, device="cuda")
    rng = torch.empty(2, dtype=torch.int32, device="cuda")
    for N in (0, 65537):
        rc = L.RecviewOutputStageNoAlloc(C.c_void_p(img.data_ptr()), 0, N, 2.0, 2, C.c_void_p(out.data_ptr()),
                                         C.c_void_p(rng.data_ptr()), None)
        assert rc == -1
    assert b"65536" in L.RecviewLastMessage()

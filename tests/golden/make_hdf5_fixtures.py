"""Writes the HDF5 fixtures of tests/test_hdf5.py with the HDF5 C library.

The files mimic APS DXchange scans as RecView reads them (gazoDoc.cpp:2229-2312,
general.cpp:1668-1683): /exchange/data (frames x rows x columns, uint16) and,
for step scans, /exchange/theta (one angle in degrees per frame).  Three files:

  dxchange_step_f32.h5         default file format (superblock 0, symbol-table
                               groups), chunked data, float32 theta, plus
                               data_dark / data_white datasets in the group
  dxchange_step_f64_latest.h5  newest file format (superblock 3, link messages
                               in version-2 object headers), contiguous data,
                               float64 theta
  dxchange_fly.h5              no theta: a fly scan (angles from the frame count)

hdf5_fixtures.json holds what was written (frame counts, theta values).
Run here with `python3 tests/golden/make_hdf5_fixtures.py`; it needs libhdf5
(HDF5_LIB, default /opt/conda/lib/libhdf5.so.103).
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.environ.get("HDF5_LIB", "/opt/conda/lib/libhdf5.so.103")

H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_LIBVER_LATEST = 2          # 1.10: H5F_LIBVER_V110 == LATEST


def main():
    L = C.CDLL(LIB)
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    hid = C.c_int64
    for f in ("H5Fcreate", "H5Gcreate2", "H5Screate_simple", "H5Dcreate2", "H5Pcreate"):
        getattr(L, f).restype = hid
    L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid, hid]
    L.H5Gcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid]
    L.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.H5Dcreate2.argtypes = [hid, C.c_char_p, hid, hid, hid, hid, hid]
    L.H5Dwrite.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    L.H5Pcreate.argtypes = [hid]
    L.H5Pset_chunk.argtypes = [hid, C.c_int, C.POINTER(C.c_uint64)]
    L.H5Pset_libver_bounds.argtypes = [hid, C.c_int, C.c_int]
    for f in ("H5Dclose", "H5Sclose", "H5Gclose", "H5Fclose", "H5Pclose"):
        getattr(L, f).argtypes = [hid]
    T_FLOAT = hid.in_dll(L, "H5T_NATIVE_FLOAT_g").value
    T_DOUBLE = hid.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    T_USHORT = hid.in_dll(L, "H5T_NATIVE_USHORT_g").value
    P_DCREATE = hid.in_dll(L, "H5P_CLS_DATASET_CREATE_ID_g").value
    P_FACCESS = hid.in_dll(L, "H5P_CLS_FILE_ACCESS_ID_g").value

    def check(v, what):
        if v < 0:
            raise RuntimeError(what + " failed")
        return v

    def dataset(group, name, arr, mtype, chunk=None):
        dims = (C.c_uint64 * arr.ndim)(*arr.shape)
        sp = check(L.H5Screate_simple(arr.ndim, dims, None), "H5Screate_simple")
        dcpl = H5P_DEFAULT
        if chunk is not None:
            dcpl = check(L.H5Pcreate(P_DCREATE), "H5Pcreate")
            ch = (C.c_uint64 * arr.ndim)(*chunk)
            check(L.H5Pset_chunk(dcpl, arr.ndim, ch), "H5Pset_chunk")
        d = check(L.H5Dcreate2(group, name.encode(), mtype, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT), "H5Dcreate2")
        a = np.ascontiguousarray(arr)
        check(L.H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), "H5Dwrite")
        L.H5Dclose(d)
        L.H5Sclose(sp)
        if chunk is not None:
            L.H5Pclose(dcpl)

    def write(fname, frames, shape, theta=None, theta_type=None, latest=False, chunked=False):
        fapl = H5P_DEFAULT
        if latest:
            fapl = check(L.H5Pcreate(P_FACCESS), "H5Pcreate")
            check(L.H5Pset_libver_bounds(fapl, H5F_LIBVER_LATEST, H5F_LIBVER_LATEST), "H5Pset_libver_bounds")
        path = os.path.join(HERE, fname)
        f = check(L.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, fapl), "H5Fcreate")
        g = check(L.H5Gcreate2(f, b"exchange", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate2")
        rng = np.random.default_rng(frames)
        data = rng.integers(0, 4096, (frames,) + shape).astype(np.uint16)
        if chunked:
            dataset(g, "data_dark", data[:2], T_USHORT, (1,) + shape)
            dataset(g, "data_white", data[:3], T_USHORT, (1,) + shape)
        dataset(g, "data", data, T_USHORT, (1,) + shape if chunked else None)
        if theta is not None:
            dataset(g, "theta", theta, theta_type)
        L.H5Gclose(g)
        L.H5Fclose(f)
        if latest:
            L.H5Pclose(fapl)

    theta32 = (np.arange(12, dtype=np.float64) * 15.25 + 0.125).astype(np.float32)
    theta64 = np.array([0.0, 19.999999, 40.000001, 60.5, 80.0, 100.25, 120.0, 140.0, 160.0, 179.75])
    write("dxchange_step_f32.h5", 12, (4, 6), theta32, T_FLOAT, chunked=True)
    write("dxchange_step_f64_latest.h5", 10, (3, 5), theta64, T_DOUBLE, latest=True)
    write("dxchange_fly.h5", 20, (2, 3))
    meta = {
        "dxchange_step_f32.h5": {"frames": 12, "theta": [float(v) for v in theta32]},
        "dxchange_step_f64_latest.h5": {"frames": 10, "theta": [float(v) for v in theta64]},
        "dxchange_fly.h5": {"frames": 20, "theta": None},
    }
    with open(os.path.join(HERE, "hdf5_fixtures.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print("wrote", ", ".join(meta))


if __name__ == "__main__":
    main()

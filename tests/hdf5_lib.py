"""Minimal reader of 1-D HDF5 datasets through the HDF5 C library (ctypes),
for checking snapshot files in tests (h5py is not installed)."""
import ctypes
import os

import numpy as np

_H5 = None


def lib():
    global _H5
    if _H5 is None:
        root = os.environ.get("MCAC_HDF5_ROOT", "/opt/conda")
        L = ctypes.CDLL(os.path.join(root, "lib", "libhdf5.so"))
        L.H5open()
        L.H5Fopen.restype = ctypes.c_int64
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_int64]
        L.H5Dopen2.restype = ctypes.c_int64
        L.H5Dopen2.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
        L.H5Dget_space.restype = ctypes.c_int64
        L.H5Dget_space.argtypes = [ctypes.c_int64]
        L.H5Sget_simple_extent_dims.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]
        L.H5Dget_type.restype = ctypes.c_int64
        L.H5Dget_type.argtypes = [ctypes.c_int64]
        L.H5Tget_size.restype = ctypes.c_size_t
        L.H5Tget_size.argtypes = [ctypes.c_int64]
        L.H5Tget_class.argtypes = [ctypes.c_int64]
        L.H5Dread.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                              ctypes.c_void_p]
        for f in ("H5Dclose", "H5Sclose", "H5Tclose", "H5Fclose"):
            getattr(L, f).argtypes = [ctypes.c_int64]
        _H5 = L
    return _H5


def read_dataset(path: str, name: str) -> np.ndarray:
    """A 1-D dataset as float64 (float class) or int64 (integer class)."""
    L = lib()
    f = L.H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
    assert f >= 0, path
    try:
        d = L.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        sp = L.H5Dget_space(d)
        dims = (ctypes.c_uint64 * 1)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        t = L.H5Dget_type(d)
        is_float = L.H5Tget_class(t) == 1  # H5T_FLOAT
        mem = ctypes.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g" if is_float else "H5T_NATIVE_LLONG_g").value
        out = np.zeros(int(dims[0]), np.float64 if is_float else np.int64)
        rc = L.H5Dread(d, mem, 0, 0, 0, out.ctypes.data)  # H5S_ALL, H5S_ALL, H5P_DEFAULT
        assert rc >= 0
        L.H5Tclose(t)
        L.H5Sclose(sp)
        L.H5Dclose(d)
        return out
    finally:
        L.H5Fclose(f)


def read_xmf(path: str):
    """(physics information dict, list of grids) as pymcac's xdmf reader sees them:
    each grid = {"Name", "GridTime": float, "Positions": dataset, <attribute>: (dataset, size, file)}."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    info = {e.get("Name"): e.get("Value") for e in root.iter("Information")}
    grids = []
    for g in root.iter("Grid"):
        if g.get("Name") == "Collection":
            continue
        rec = {"Name": g.get("Name"), "GridTime": float(g.find("Time").get("Value")),
               "Positions": list(g.find("Geometry"))[0].text.split(":")[-1]}
        for a in g.findall("Attribute"):
            item = list(a)[0]
            rec[a.get("Name")] = (item.text.split(":")[-1], int(item.get("Dimensions")), item.text.split(":")[0])
        grids.append(rec)
    return info, grids

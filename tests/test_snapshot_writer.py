"""Snapshot files in the reference's XDMF/HDF5 layout (src/io/writer.cpp,
format.cpp, sphere_list.cpp), read back the way pymcac's readers do."""
import math
import os

import numpy as np

import hdf5_lib as H
import mcac_amd


def test_snapshot_layout_and_values(tmp_path):
    n, grids, per_file = 37, 5, 2
    prefix = str(tmp_path / "Spheres")
    assert mcac_amd.load_sim().mcac_write_test_snapshots(prefix.encode(), n, grids, per_file) == 0
    width = math.ceil(math.log10(n)) + 4  # format.cpp:60-65
    names = sorted(os.listdir(tmp_path))
    expect = [f"Spheres_{k:0{width}d}.{ext}" for k in range(3) for ext in ("h5", "xmf")]
    assert names == sorted(expect)
    seen = 0
    for k in range(3):
        xmf = tmp_path / f"Spheres_{k:0{width}d}.xmf"
        info, gl = H.read_xmf(str(xmf))
        assert info["Copyright"] == "Produced by MCAC" and info["L"] == "1e-06" and info["N []"] == str(n)
        assert len(gl) == (2 if k < 2 else 1)
        for g in gl:
            gidx = seen
            seen += 1
            assert g["Name"] == "Spheres" and g["GridTime"] == 0.5 * gidx
            h5 = str(tmp_path / f"Spheres_{k:0{width}d}.h5")
            pos = H.read_dataset(h5, g["Positions"]).reshape(n, 3)
            kk = np.arange(n)
            assert np.array_equal(pos[:, 0], (gidx + kk) / 8.0)
            assert np.array_equal(pos[:, 1], 2 * kk / 8.0) and np.array_equal(pos[:, 2], 3 * kk / 8.0)
            assert [a for a in g if a not in ("Name", "GridTime", "Positions")] == ["Time", "electric_charge", "Radius",
                                                                                  "Label"]
            assert g["Radius"][1] == n and g["Label"][2] == f"Spheres_{k:0{width}d}.h5"
            assert np.array_equal(H.read_dataset(h5, g["Radius"][0]), kk / 4.0)
            assert np.array_equal(H.read_dataset(h5, g["Label"][0]), kk)
            assert np.array_equal(H.read_dataset(h5, g["electric_charge"][0]), np.zeros(n))
            assert H.read_dataset(h5, g["Time"][0])[0] == 0.5 * gidx
    assert seen == grids

"""The C-ABI libraries load and export every symbol include/*.h declares (no GPU needed)."""
import ctypes
import pathlib
import re

import mcac_amd

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared(header):
    text = (ROOT / "include" / header).read_text()
    return set(re.findall(r"^\s*(?:int|int64_t|double|const char \*|mcac_box \*)\s*\*?(mcac_\w+)\(", text, re.M))


def test_headers_match_binding_lists():
    assert declared("mcac_gpu.h") == set(mcac_amd.GPU_SYMBOLS)
    assert declared("mcac_sim.h") == set(mcac_amd.SIM_SYMBOLS)


def test_gpu_library_exports():
    lib = mcac_amd.load_gpu()
    for s in mcac_amd.GPU_SYMBOLS:
        assert hasattr(lib, s), s
    assert b"gfx950" in lib.mcac_version()


def test_sim_library_exports():
    lib = mcac_amd.load_sim()
    for s in mcac_amd.SIM_SYMBOLS:
        assert hasattr(lib, s), s


def test_record_layouts():
    assert mcac_amd.STEP_DTYPE.itemsize == 112
    assert mcac_amd.MERGE_DTYPE.itemsize == 176
    assert mcac_amd.CONTACT_DTYPE.itemsize == 80
    assert mcac_amd.MERGE_RESULT_DTYPE.itemsize == 144
    assert mcac_amd.Box.PARAMS.itemsize == 64

import os
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (str(ROOT), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

# every cell-list build whose starts the host counts from its mirrors is also
# read back from the device and compared (a mismatch is a status error)
os.environ.setdefault("MCAC_VERIFY_CELLS", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")
    config.addinivalue_line("markers", "slow: longer parity runs")

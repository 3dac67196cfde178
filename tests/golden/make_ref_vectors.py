"""Regenerate tests/golden/ref_vectors.json from the reference's own sources.

Builds oracle/_ref (reference src/tools/tools.cpp and src/verlet/verlet.cpp
compiled as-is by oracle/Makefile.ref) and runs its probe, which writes the
glibc rand() stream, random_direction(), inverf()/inverfc() and Verlet
get_neighborhood() orders as JSON (doubles as C99 hex floats).

Run in the build container (needs /root/reference):  python tests/golden/make_ref_vectors.py
"""
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[2]


def main() -> int:
    subprocess.run(["make", "-s", "-f", str(ROOT / "oracle" / "Makefile.ref")], check=True)
    out = subprocess.run([str(ROOT / "oracle" / "_ref" / "ref_probe")], check=True,
                         capture_output=True, text=True).stdout
    dst = ROOT / "tests" / "golden" / "ref_vectors.json"
    dst.write_text(out)
    print(f"wrote {dst} ({len(out)} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())

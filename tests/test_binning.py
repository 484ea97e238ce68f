"""The cell index of the binning, the tile sort and the list build (spp_amd/csrc/binning.hpp), on
the CPU: the host build of the same code the device runs.  It replaces the division
floor((x + L) nc / 2L) by a product with the rounded reciprocal and falls back to the division
within a few ulps of a cell edge; it must equal the division everywhere, which is what keeps the
cells -- and so the neighbour sets -- those of the reference's Cell binning (shared/cell.h)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cell_index_equals_division(tmp_path):
    exe = str(tmp_path / "cell_index_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "cpp", "cell_index_check.cpp"), "-o", exe], check=True)
    n, bad = (int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split())
    assert n == 6_000_000 and bad == 0

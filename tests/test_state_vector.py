"""The facade's State (the reference's State_Hyper_Vector, shared/state-hyper-vector.h) and the
formation helpers, on the CPU.

* State: a difference of two nearby states is small also across the periodic edges (minimum
  image), adding it back restores the state (r mod the box, theta mod 2 pi, unwrapped positions
  exactly); 9.9 and -9.9 in a box of half width 10 differ by -0.2; Square wraps theta into
  (-pi, pi]; Unit has magnitude 1; Max_Index finds the largest |theta|.
* Random_Formation with the sigma inset keeps every position within L - sigma; Polar_Formation
  has the same positions and every heading 0; sigma = 0 reproduces spp_random_state.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import spp_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("state") / "state_vector")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "state_vector.cpp"), "-o", exe], check=True)
    lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    return {l.split()[0]: l.split()[1:] for l in lines if l}


def test_deviation_and_round_trip(out):
    assert float(out["max_dev"][0]) < 2e-3
    er, eu = (float(v) for v in out["roundtrip"])
    assert er < 1e-12 and eu < 1e-12


def test_minimum_image_square_unit_max_index(out):
    assert abs(float(out["edge"][0]) - (-0.2)) < 1e-12
    assert abs(float(out["square"][0]) - (math.pi / 2) ** 2) < 1e-12
    mag, diff = (float(v) for v in out["unit"])
    assert abs(mag - 1) < 1e-14 and diff < 1e-18
    assert int(out["maxindex"][0]) == 1


def test_random_formation_inset_and_polar_formation(engine_mod):
    x, y, th = S.random_formation(5, 4000, 12.0, 7.0, sigma=1.5)
    assert np.all(np.abs(x) <= 10.5) and np.all(np.abs(y) <= 5.5)
    assert x.max() > 10.3 and y.min() < -5.3  # fills the inset box
    assert np.all((th >= 0) & (th < 2 * np.pi))
    x0, y0, th0 = S.random_formation(5, 4000, 12.0, 7.0, sigma=0.0)
    xr, yr, thr = S.random_state(5, 4000, 12.0, 7.0)
    assert np.array_equal(x0, xr) and np.array_equal(y0, yr) and np.array_equal(th0, thr)
    xp, yp, tp = S.polar_formation(5, 4000, 12.0, 7.0)
    assert np.array_equal(xp, xr) and np.array_equal(yp, yr) and np.all(tp == 0.0)
    with pytest.raises(S.SppError):
        S.random_formation(5, 10, 12.0, 7.0, sigma=7.5)

"""The heading sine/cosine of the stage kernels and of the noise transform
(spp_amd/csrc/trig.hpp), on the CPU: the host build of the same code the device runs.

Pinned against extended-precision sinl/cosl (x87, 64-bit significand): at most 1.5 ulp over
[-4, 4] and near multiples of pi/2 (where the reduction cancels), absolute error below 1.2e-16,
and the exact points sin 0 = 0, cos 0 = 1, sin(pi/2) = 1, cos(+-pi) = -1.
"""
import math
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("trig") / "trig_accuracy")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "cpp", "trig_accuracy.cpp"), "-o", exe], check=True)
    return subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")


def test_error_bound(out):
    es, ec, ea = (float(v) for v in out[0].split())
    assert es <= 1.5 and ec <= 1.5 and ea < 1.2e-16


def test_exact_points(out):
    vals = {float(l.split()[0]): tuple(float(v) for v in l.split()[1:]) for l in out[1:] if l}
    assert vals[0.0] == (0.0, 1.0)
    assert vals[math.pi / 2][0] == 1.0 and abs(vals[math.pi / 2][1]) < 1e-16
    assert vals[math.pi][1] == -1.0 and vals[-math.pi][1] == -1.0
    for x in (0.5, -2.5):
        assert vals[x] == pytest.approx((math.sin(x), math.cos(x)), rel=3e-16, abs=1e-17)


@pytest.fixture(scope="module")
def sin_diff_out(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("trig") / "sin_diff_accuracy")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "cpp", "sin_diff_accuracy.cpp"), "-o", exe], check=True)
    return subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")


def test_sin_diff_relative_error(sin_diff_out):
    """sin(a - b) of the alignment torque stays within a few ulps *relative* for uniform, nearly
    aligned (|a - b| down to 1e-14), nearly anti-aligned and wrap-aligned heading pairs -- the
    product form s_j c_i - c_j s_i is ~1e-16 absolute, i.e. 1e-2 relative at |a - b| = 1e-14."""
    uni, aligned, anti, wrapped = (float(v) for v in sin_diff_out[0].split())
    assert uni <= 2.5
    assert aligned <= 1.5 and anti <= 1.5 and wrapped <= 1.5


def test_sin_diff_exact_values(sin_diff_out):
    z, t, u, one = (float(v) for v in sin_diff_out[1].split())
    assert z == 0.0 and t == 1e-12 and u == 1e-300
    assert abs(one - 1.0) <= 2.3e-16


@pytest.fixture(scope="module")
def log_out(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("trig") / "log_accuracy")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "cpp", "log_accuracy.cpp"), "-o", exe], check=True)
    return subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")


def test_log_error_bound(log_out):
    """The noise draw's logarithm: within one ulp of extended-precision logl on the draw's own
    arguments 1 - u, next to 1 and to the reduction's boundaries, and over [2^-60, 2^10]."""
    noise, edges, spread = (float(v) for v in log_out[0].split())
    assert noise < 1.0 and edges < 1.0 and spread < 1.0


def test_log_exact_values(log_out):
    one, two, half, tiny = (float(v) for v in log_out[1].split())
    assert one == 0.0
    assert two == pytest.approx(math.log(2.0), rel=2.3e-16) and half == -two
    assert tiny == pytest.approx(-53 * math.log(2.0), rel=2.3e-16)

"""Vesicle observables of the Box facade (spp_amd/csrc/membrane.hpp; reference
Box::Compute_All_Variables, serial/beadbox.h:433-496) on the CPU, pinned by closed forms:

* a regular ring of radius R: gyration radius R, asphericity 0, centre of mass at its centre;
* an ellipse with semi-axes a, b at any tilt: Rg^2 = (a^2 + b^2) / 2 and asphericity
  ((a^2 - b^2) / (a^2 + b^2))^2 (tilt invariance pins the off-diagonal term);
* the ring turned rigidly by phi about its centre and the whole vesicle shifted by d over a time
  dt: centre-of-mass velocity d / dt and angular frequency sin(phi) / dt; swimmers (their centre
  off the ring's) turned by psi about the membrane's centre: swimmer angular frequency
  sin(psi) / dt, whatever phi and d are.
"""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("membrane") / "membrane_vars")
    subprocess.run(["g++", "-O1", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "membrane_vars.cpp"), "-o", exe],
                   check=True)
    return exe


def run(prog, x, y, xo, yo, nm, dt):
    txt = "%d %d %.17g\n" % (nm, len(x) - nm, dt)
    txt += "".join("%.17g %.17g %.17g %.17g\n" % r for r in zip(x, y, xo, yo))
    out = subprocess.run([prog], input=txt, capture_output=True, text=True, check=True).stdout.split("\n")
    vals = {}
    for line in out[:15]:
        k, v = line.split()
        vals[k] = float(v)
    return vals, out[15:17]


def rot(x, y, a):
    c, s = math.cos(a), math.sin(a)
    return c * x - s * y, s * x + c * y


def test_ring_and_ellipse_shape(prog):
    nm = 97
    a = 2 * np.pi * np.arange(nm) / nm
    R, c = 8.5, (3.0, -2.0)
    x, y = c[0] + R * np.cos(a), c[1] + R * np.sin(a)
    v, _ = run(prog, x, y, x, y, nm, 0.1)
    assert v["mb_Rg"] == pytest.approx(R, rel=1e-12)
    assert v["mb_Delta"] < 1e-20
    assert (v["mb_x_cm"], v["mb_y_cm"]) == pytest.approx(c, abs=1e-12)
    assert v["mb_omega"] == 0.0 and v["mb_vx_cm"] == 0.0
    A, B = 9.0, 5.0
    for tilt in (0.0, 0.4, 1.3):
        ex, ey = rot(A * np.cos(a), B * np.sin(a), tilt)
        v, _ = run(prog, ex + 1.0, ey, ex + 1.0, ey, nm, 0.1)
        assert v["mb_Rg"] ** 2 == pytest.approx((A * A + B * B) / 2, rel=1e-12)
        assert v["mb_Delta"] == pytest.approx(((A * A - B * B) / (A * A + B * B)) ** 2, rel=1e-10)
        q = np.array([[v["mb_xx"], v["mb_xy"]], [v["mb_xy"], v["mb_yy"]]])
        assert np.linalg.eigvalsh(q) == pytest.approx([B * B / 2, A * A / 2], rel=1e-10)


@pytest.mark.parametrize("phi,psi", [(0.01, -0.03), (-0.2, 0.0), (0.0, 0.5)])
def test_rigid_rotation_and_translation(prog, phi, psi):
    nm, ns = 64, 200
    rng = np.random.default_rng(3)
    a = 2 * np.pi * np.arange(nm) / nm
    c = np.array([-4.0, 6.0])
    ux, uy = 10 * np.cos(a), 10 * np.sin(a)
    r = 9 * np.sqrt(rng.uniform(0, 1, ns))
    t = rng.uniform(0, 2 * np.pi, ns)
    wx, wy = r * np.cos(t), r * np.sin(t)
    wx += 2.5  # swimmers off the ring's centre: the rotation is still about the ring's centre
    d, dt = np.array([0.3, -0.7]), 0.05
    xo = np.concatenate([c[0] + ux, c[0] + wx])
    yo = np.concatenate([c[1] + uy, c[1] + wy])
    mx, my = rot(ux, uy, phi)
    sx, sy = rot(wx, wy, psi)
    x = np.concatenate([c[0] + d[0] + mx, c[0] + d[0] + sx])
    y = np.concatenate([c[1] + d[1] + my, c[1] + d[1] + sy])
    v, (head, row) = run(prog, x, y, xo, yo, nm, dt)
    assert (v["mb_vx_cm"], v["mb_vy_cm"]) == pytest.approx(tuple(d / dt), rel=1e-10)
    assert v["mb_omega"] == pytest.approx(math.sin(phi) / dt, rel=1e-9, abs=1e-9)
    assert v["sw_omega"] == pytest.approx(math.sin(psi) / dt, rel=1e-9, abs=1e-9)
    swc = rot(wx.mean(), wy.mean(), psi)
    assert (v["sw_x_cm"], v["sw_y_cm"]) == pytest.approx((c[0] + d[0] + swc[0], c[1] + d[1] + swc[1]), abs=1e-9)
    # the quantities file: header then time + 12 values (beadbox.h:502-504)
    assert head.split("\t")[:3] == ["#", "time", "x_cm"] and len(head.split("\t")) == 14
    cols = [float(s) for s in row.split("\t")]
    assert len(cols) == 13 and cols[0] == 1.0
    assert cols[6] == pytest.approx(v["mb_omega"], rel=1e-5) and cols[12] == pytest.approx(v["sw_omega"], rel=1e-5,
                                                                                               abs=1e-6)

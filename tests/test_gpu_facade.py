"""The C++ Box facade (spp_amd/csrc/box.hpp) driving the GPU engine: the reference's
data-gathering loop, its binary trajectory and resume from the last healthy frame."""
import os
import struct
import subprocess

import numpy as np
import pytest

import spp_amd as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "spp_amd", "box_demo")


def read_frames(path):
    data = open(path, "rb").read()
    out, off = [], 0
    while off + 36 <= len(data):
        t, Lx, Ly, nb, Ns, Nm = struct.unpack_from("<dddiii", data, off)
        off += 36
        m = np.frombuffer(data, "<f4", 2 * Nm, off).reshape(Nm, 2)
        off += 8 * Nm
        s = np.frombuffer(data, "<f4", 3 * Ns, off).reshape(Ns, 3)
        off += 12 * Ns
        out.append((t, Lx, Ly, nb, s))
    assert off == len(data)
    return out


def params(n, L):
    p = S.default_params(S.MODEL_REPULSIVE)
    return p.update(n=n, Lx=L, Ly=L, dt=0.005, v0=1.0, R_align=1.0, repulsion=S.REP_HARMONIC,
                    A_rep=10.0, rc_rep=1.0, D_rot=0.05, r_list=1.3, seed=2)


def test_box_demo_trajectory_and_resume(tmp_path):
    assert os.path.exists(DEMO), "build() compiles spp_amd/box_demo"
    n, L = 4096, 32.0
    traj, pol = str(tmp_path / "traj.bin"), str(tmp_path / "pol.dat")
    subprocess.run([DEMO, traj, pol, str(n), str(L), "200", "20", "2"], check=True)
    fr = read_frames(traj)
    # initial frame + one every 2 chunks of 20 steps
    assert [round(f[0], 9) for f in fr] == [0.0, 0.2, 0.4, 0.6, 0.8, 1.0]
    assert fr[0][1] == L and fr[0][3] == 1 and fr[0][4].shape == (n, 3)
    # the same run through the Python binding: positions are r_original in float32
    e = S.Engine(params(n, L))
    x, y, th = S.random_state(7, n, L, L)
    e.set_state(x, y, th)
    e.multi_step(200)
    st = e.get_state()
    last = fr[-1][4]
    assert np.abs(last[:, 0] - st["xu"].astype(np.float32)).max() < 1e-3
    assert np.abs(last[:, 2] - st["theta"].astype(np.float32)).max() < 1e-3
    lines = open(pol).read().split("\n")
    assert len([l for l in lines if l]) == 10
    # resume: continues from the last frame's time and appends
    subprocess.run([DEMO, traj, pol, str(n), str(L), "80", "20", "2", "1"], check=True)
    fr2 = read_frames(traj)
    assert len(fr2) == len(fr) + 2
    assert round(fr2[-1][0], 9) == 1.4

"""The C++ Box facade (spp_amd/csrc/box.hpp) driving the GPU engine: the reference's
data-gathering loop, its binary trajectory and resume from the last healthy frame."""
import math
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import spp_amd as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "spp_amd", "box_demo")
VESICLE = os.path.join(ROOT, "spp_amd", "vesicle_demo")


def read_frames(path, with_theta=True, membrane=False):
    data = open(path, "rb").read()
    out, off = [], 0
    k = 3 if with_theta else 2
    while off + 36 <= len(data):
        t, Lx, Ly, nb, Ns, Nm = struct.unpack_from("<dddiii", data, off)
        off += 36
        m = np.frombuffer(data, "<f4", 2 * Nm, off).reshape(Nm, 2)
        off += 8 * Nm
        s = np.frombuffer(data, "<f4", k * Ns, off).reshape(Ns, k)
        off += 4 * k * Ns
        out.append((t, Lx, Ly, nb, s, m) if membrane else (t, Lx, Ly, nb, s))
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
    run = subprocess.run([DEMO, traj, pol, str(n), str(L), "200", "20", "2"], check=True, capture_output=True, text=True)
    # operator>> reads every frame back; the last one holds the final state to float32 rounding
    m = re.search(r"read back (\d+) frames, last t=([0-9.]+), max error (\S+)", run.stdout)
    assert m and int(m.group(1)) == 6 and abs(float(m.group(2)) - 1.0) < 1e-9 and float(m.group(3)) < 1e-5, run.stdout
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


def vesicle_params(nm, ns, F0):
    """The parameters vesicle_demo.cpp sets: touching unit beads on the ring, harmonic repulsion,
    bonds k = 100, bending 10, D_r = 1/8, RK2 dt = 1e-3."""
    R = 0.5 / math.sin(math.pi / nm)
    p = S.default_params(S.MODEL_ABP)
    p.update(n=nm + ns, n_membrane=nm, Lx=R + 6, Ly=R + 6, integrator=S.RK2, dt=1e-3, repulsion=S.REP_HARMONIC,
             A_rep=10.0, sigma_rep=1.0, rc_rep=1.0, r_list=1.3, D_rot=0.125, v0=F0, mob_par=0.5, mob_perp=0.5,
             mob_rot=1.0, k_bond=100.0, l_bond=1.0, kappa_bend=10.0, seed=5)
    return p, R


@pytest.mark.parametrize("F0", [2.0, 0.0])
def test_vesicle_demo_frames_observables_and_resume(tmp_path, F0):
    """serial/membrane.cpp's loop: frames carry the heading only for self-propelled swimmers
    (particle.h:1038-1045); the quantities file's gyration radius and asphericity agree with the
    engine's state computed here with numpy; a resumed run appends from the last frame."""
    assert os.path.exists(VESICLE), "build() compiles spp_amd/vesicle_demo"
    nm, ns = 128, 512
    traj, quant = str(tmp_path / "ves.bin"), str(tmp_path / "q.dat")
    subprocess.run([VESICLE, traj, quant, str(nm), str(ns), str(F0), "200", "20", "2", "1"], check=True)
    th = F0 != 0
    assert os.path.getsize(traj) == 6 * (36 + 8 * nm + (12 if th else 8) * ns)
    fr = read_frames(traj, with_theta=th, membrane=True)
    assert [round(f[0], 9) for f in fr] == [0.0, 0.04, 0.08, 0.12, 0.16, 0.2]
    p, R = vesicle_params(nm, ns, F0)
    e = S.Engine(p)
    mx, my = S.ring_membrane(nm, 1.0, R)
    sx, sy, sth = S.disk_state(11, ns, R - 1.0)
    e.set_state(np.concatenate([mx, sx]), np.concatenate([my, sy]), np.concatenate([np.zeros(nm), sth]))
    e.multi_step(200)
    st = e.get_state()
    last = fr[-1]
    assert np.abs(last[5][:, 0] - st["xu"][:nm].astype(np.float32)).max() < 1e-3
    assert np.abs(last[4][:, 1] - st["yu"][nm:].astype(np.float32)).max() < 1e-3
    if th:
        assert np.abs(last[4][:, 2] - st["theta"][nm:].astype(np.float32)).max() < 1e-3
    rows = [l for l in open(quant).read().split("\n") if l]
    assert rows[0].startswith("#\ttime") and len(rows) == 11
    cols = [float(c) for c in rows[-1].split("\t")]
    bx, by = st["xu"][:nm], st["yu"][:nm]
    assert cols[0] == pytest.approx(0.2)
    assert cols[7] == pytest.approx(math.sqrt(bx.var() + by.var()), rel=1e-4)
    ev = np.linalg.eigvalsh(np.cov(np.stack([bx, by]), bias=True))
    assert cols[8] == pytest.approx(((ev[1] - ev[0]) / (ev[1] + ev[0])) ** 2, rel=1e-3, abs=1e-8)
    subprocess.run([VESICLE, traj, quant, str(nm), str(ns), str(F0), "40", "20", "1", "1", "1"], check=True)
    fr2 = read_frames(traj, with_theta=th, membrane=True)
    assert len(fr2) == len(fr) + 2 and round(fr2[-1][0], 9) == 0.24
    rows2 = [l for l in open(quant).read().split("\n") if l]
    assert len(rows2) == len(rows) + 2 and not rows2[-1].startswith("#")

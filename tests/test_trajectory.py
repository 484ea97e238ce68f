"""Trajectory frames in the reference's binary format (CPU).

The byte layout is encoded here independently with struct, from the reference writer
(serial/box.h:549-568) and reader (analyze/read.h:215-274), and fed to the C++ codec used by the
Box facade."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def frame_bytes(t, L, Ns, Nm, rng, with_theta=True):
    b = struct.pack("<dddiii", t, L, L, 1, Ns, Nm)
    m = rng.uniform(-L, L, (Nm, 2)).astype("<f4")
    s = rng.uniform(-L, L, (Ns, 3 if with_theta else 2)).astype("<f4")
    return b + m.tobytes() + s.tobytes(), m, s


@pytest.fixture(scope="module")
def codec(tmp_path_factory):
    d = tmp_path_factory.mktemp("codec")
    exe = str(d / "frame_roundtrip")
    subprocess.run(["g++", "-O1", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "frame_roundtrip.cpp"), "-o", exe],
                   check=True)
    return exe


def run(codec, tmp_path, data, with_theta=True):
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    src.write_bytes(data)
    out = subprocess.run([codec, str(src), "1" if with_theta else "0", str(dst)], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    return [l.split() for l in out if l], dst.read_bytes()


def test_frame_size_and_roundtrip(codec, tmp_path):
    rng = np.random.default_rng(0)
    f1, m1, s1 = frame_bytes(0.5, 32.0, 100, 16, rng)
    f2, _, _ = frame_bytes(1.0, 32.0, 100, 16, rng)
    assert len(f1) == 36 + 8 * 16 + 12 * 100
    lines, out = run(codec, tmp_path, f1 + f2)
    assert out == f1 + f2
    assert [float(l[0]) for l in lines] == [0.5, 1.0]
    assert int(lines[0][1]) == 100 and int(lines[0][2]) == 16
    assert float(lines[0][3]) == pytest.approx(float(s1[0, 0]))


def test_unhealthy_and_truncated_frames_dropped(codec, tmp_path):
    """fix-file semantics (analyze/fix-file.cpp, read.h:230-271): t outside [0, 1e6] marks a frame
    unhealthy; a frame cut short by a killed run ends the file."""
    rng = np.random.default_rng(1)
    good, _, _ = frame_bytes(2.0, 8.0, 10, 0, rng)
    bad, _, _ = frame_bytes(-1.0, 8.0, 10, 0, rng)
    tail, _, _ = frame_bytes(3.0, 8.0, 10, 0, rng)
    lines, out = run(codec, tmp_path, good + bad + good + tail[:-7])
    assert out == good + good
    assert len(lines) == 2


def test_frames_without_theta(codec, tmp_path):
    """ActiveBrownianChain with F0 == 0 writes no heading (particle.h:1038-1039)."""
    rng = np.random.default_rng(2)
    f, _, _ = frame_bytes(1.5, 8.0, 7, 3, rng, with_theta=False)
    assert len(f) == 36 + 8 * 3 + 8 * 7
    lines, out = run(codec, tmp_path, f + f, with_theta=False)
    assert out == f + f

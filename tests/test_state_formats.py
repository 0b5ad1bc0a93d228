"""State formats (ow3d/io.py) against the reference's own files and writer:

* bottom file: the reference example SF_Shoaling_BottomFile (tests/golden/ref_SF_Shoaling_BottomFile.txt,
  137 x 1 nodes, annotated first data line) and the ghost-node extension of SetupInitialConditions.f90:92-108;
* hot-start file: the reference example SFWaveAndFollowingCurrent.init (641 x 1 nodes);
* EP_#####.bin: byte-identical to the file the reference's StoreData (IO/StoreData.f90, compiled into
  oracle/_ref) writes for the same arrays, and read back exactly.
"""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _io():
    from ow3d import io
    return io


def test_reference_bottom_file_reads_and_round_trips(tmp_path):
    io = _io()
    b = io.read_bottom(os.path.join(HERE, "ref_SF_Shoaling_BottomFile.txt"), 137, 1)
    assert b["h"].shape == (1, 139)
    assert b["h"][0, 1] == 50.0 and b["h"][0, 0] == b["h"][0, 1] and b["h"][0, -1] == b["h"][0, -2]
    assert b["hx"][0, 2] == -7.316045e-28
    # smooth beach from 50 m to 15 m
    assert abs(b["h"][0, -2] - 15.0) < 1e-6 and np.all(np.diff(b["h"][0, 1:-1]) <= 1e-12)
    p = tmp_path / "bottom"
    io.write_bottom(p, b, 137, 1)
    b2 = io.read_bottom(p, 137, 1)
    for k in b:
        assert np.array_equal(b[k], b2[k]), k


def test_reference_hotstart_file_reads_and_round_trips(tmp_path):
    io = _io()
    E, P, (lx, ly, t0) = io.read_hotstart(os.path.join(HERE, "ref_SFWaveAndFollowingCurrent.init.txt"), 641, 1)
    assert (lx, ly, t0) == (10.0, 0.0, 0.0)
    assert E.shape == (1, 643) and P[0, 1] == 0.0 and P[0, 2] == 1.5625e-03
    with pytest.raises(ValueError):
        io.read_hotstart(os.path.join(HERE, "ref_SFWaveAndFollowingCurrent.init.txt"), 640, 1)
    rng = np.random.default_rng(0)
    E3 = rng.standard_normal((19, 21)); P3 = rng.standard_normal((19, 21))
    io.write_hotstart(tmp_path / "OceanWave3D.end", E3, P3, 4.0, 3.5, 12.25)
    E4, P4, info = io.read_hotstart(tmp_path / "OceanWave3D.end", 19, 17)
    assert info == (4.0, 3.5, 12.25)
    assert np.array_equal(E4[1:-1, 1:-1], E3[1:-1, 1:-1]) and np.array_equal(P4[1:-1, 1:-1], P3[1:-1, 1:-1])


def test_ep_file_round_trip(tmp_path):
    io = _io()
    rng = np.random.default_rng(1)
    x, y, E, P = (rng.standard_normal((7, 9)) for _ in range(4))
    io.write_ep(tmp_path / "EP_00001.bin", x, y, E, P)
    out = io.read_ep(tmp_path / "EP_00001.bin")
    for a, b in zip(out, (x, y, E, P)):
        assert np.array_equal(a, b)


@pytest.mark.skipif(not (__import__("oracle").ref_available() or os.path.isdir("/root/reference/src")),
                    reason="reference tree absent: oracle/_ref cannot be built")
def test_ep_file_matches_reference_storedata(tmp_path):
    import oracle
    from ow3d import problems
    io = _io()
    nx, ny, L = 11, 9, 5.0
    bath = problems.flat_bottom(nx, ny, L, L)
    R = oracle.RefTank(nx, ny, 5, L, L, 2, bath=bath)
    rng = np.random.default_rng(2)
    E = rng.standard_normal(R.shape2); P = rng.standard_normal(R.shape2)
    ref = R.store_ep(str(tmp_path), E, P, nr=7)
    x, y, _ = R.grid()
    X, Y = np.meshgrid(x, y)
    mine = tmp_path / "mine.bin"
    io.write_ep(mine, X, Y, E, P)
    assert open(ref, "rb").read() == open(mine, "rb").read()
    xr, yr, Er, Pr = io.read_ep(ref)
    assert np.array_equal(Er, E) and np.array_equal(Pr, P) and np.array_equal(xr, X) and np.array_equal(yr, Y)

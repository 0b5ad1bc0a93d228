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


def _kinematics_case(nx, ny, nz, alpha, nonlinear=1):
    from ow3d import problems
    L = 6.0
    bath = problems.gaussian_shoal(nx, ny, L, L)
    x, y = problems.coords(nx, ny, L, L)
    X, Y = np.meshgrid(x, y)
    rng = np.random.default_rng(nx + nz)
    E = 0.05 * np.cos(2 * np.pi * X / L) * (np.cos(2 * np.pi * Y / L) if ny > 1 else 1.0) + 0.002 * rng.standard_normal(X.shape)
    P = 0.1 * rng.standard_normal(X.shape)
    nzg = nz + 1
    phi = rng.standard_normal((len(y), len(x), nzg))
    return bath, X, Y, E, P, phi


@pytest.mark.skipif(not (__import__("oracle").ref_available() or os.path.isdir("/root/reference/src")),
                    reason="reference tree absent: oracle/_ref cannot be built")
@pytest.mark.parametrize("nx,ny,nz,alpha,nonlinear", [(11, 9, 6, 2, 1), (17, 1, 8, 3, 1), (9, 9, 6, 2, 0)])
def test_kinematics_files_match_reference_storekinematicdata(tmp_path, oracle_mod, nx, ny, nz, alpha, nonlinear):
    """Kinematics_01_00000.bin (grid) and Kinematics_01_00001.bin (eta, slopes, phi, u, v, w and the
    gradients) written by ow3d/io.py from the oracle's fields are byte-identical to what the reference's
    StoreKinematicData (formattype 21, compiled into oracle/_ref) writes for the same state; this also pins
    the oracle's velocities and velocity gradients to the reference's bit for bit."""
    import oracle
    from conftest import oracle_params
    io = _io()
    bath, X, Y, E, P, phi = _kinematics_case(nx, ny, nz, alpha, nonlinear)
    R = oracle.RefTank(nx, ny, nz, 6.0, 6.0, alpha, bath=bath)
    Er, Pr, d = R.surface(E, P)
    win = (2, nx - 1, 2, 1, max(ny - 2, 1), 1, 1, 10, 1)
    rdir = tmp_path / "ref"; rdir.mkdir()
    ref0 = R.store_kinematics(str(rdir), phi, win, 0.05, 0, nonlinear)
    ref1 = R.store_kinematics(str(rdir), phi, win, 0.05, 1, nonlinear)
    p = oracle_params(nx, ny, nz, alpha=alpha, beta=alpha, gamma=alpha, lx=6.0, ly=6.0, precond=0, nonlinear=nonlinear)
    T = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    Eo, Po, do = T.surface(E, P)
    U, V, W = T.kinematics(phi, Eo)
    G = T.kinematics_gradients(phi, Eo)
    x, y, z = T.grid()
    mdir = tmp_path / "mine"; mdir.mkdir()
    grid = dict(x=X, y=Y, h=bath["h"], hx=bath["hx"], hy=bath["hy"], z=z)
    m0 = io.write_kinematics(str(mdir), 1, 0, win, 0.05, grid=grid)
    m1 = io.write_kinematics(str(mdir), 1, 1, win, 0.05,
                             fields=dict(E=Eo, Ex=do["Ex"], Ey=do["Ey"], phi=phi, U=U, V=V, W=W, G=G))
    assert os.path.basename(m0) == os.path.basename(ref0) and os.path.basename(m1) == os.path.basename(ref1)
    assert open(m0, "rb").read() == open(ref0, "rb").read()
    a, b = open(m1, "rb").read(), open(ref1, "rb").read()
    assert len(a) == len(b)
    assert a == b


def _golden_kinematics():
    with np.load(os.path.join(HERE, "kinematics_ref.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _records(raw):
    out, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw[pos:pos + 4], dtype=np.int32)[0])
        out.append(raw[pos + 4:pos + 4 + n])
        assert np.frombuffer(raw[pos + 4 + n:pos + 8 + n], dtype=np.int32)[0] == n
        pos += 8 + n
    return out


def test_kinematics_files_match_the_reference_golden_files(tmp_path, oracle_mod):
    """The reference-written pair in tests/golden/kinematics_ref.npz (make_golden.py), reproduced byte for
    byte from the oracle's fields and ow3d/io.py on machines without the reference tree."""
    from conftest import oracle_params
    io = _io()
    G = _golden_kinematics()
    nx, ny, nz, a, L = int(G["nx"]), int(G["ny"]), int(G["nz"]), int(G["alpha"]), float(G["L"])
    p = oracle_params(nx, ny, nz, alpha=a, beta=a, gamma=a, lx=L, ly=L, precond=0)
    T = oracle_mod.OracleTank(p, G["h"], G["hx"], G["hy"], G["hxx"], G["hyy"])
    Eo, Po, do = T.surface(G["E_in"], G["P_in"])
    U, V, W = T.kinematics(G["phi"], Eo)
    Gr = T.kinematics_gradients(G["phi"], Eo)
    x, y, z = T.grid()
    X, Y = np.meshgrid(x, y)
    m0 = io.write_kinematics(str(tmp_path), 1, 0, G["win"], float(G["dt"]),
                             grid=dict(x=X, y=Y, h=G["h"], hx=G["hx"], hy=G["hy"], z=z))
    m1 = io.write_kinematics(str(tmp_path), 1, 1, G["win"], float(G["dt"]),
                             fields=dict(E=Eo, Ex=do["Ex"], Ey=do["Ey"], phi=G["phi"], U=U, V=V, W=W, G=Gr))
    assert open(m0, "rb").read() == G["file0"].tobytes()
    assert open(m1, "rb").read() == G["file1"].tobytes()


@pytest.mark.gpu
def test_gpu_kinematics_file_matches_the_reference_golden_file(tmp_path):
    """The same state on the GPU (ow3d.Tank.store_kinematics): every record of the data file within 1e-12
    (relative to the record's largest value) of the reference's; the grid file byte for byte."""
    import ow3d
    G = _golden_kinematics()
    nx, ny, nz, a, L = int(G["nx"]), int(G["ny"]), int(G["nz"]), int(G["alpha"]), float(G["L"])
    T = ow3d.Tank(dict(nx=nx, ny=ny, nz=nz, alpha=a, beta=a, gamma=a, lx=L, ly=L, precond=0),
                  G["h"], G["hx"], G["hy"], G["hxx"], G["hyy"])
    T.set_state(G["E_in"], G["P_in"], None, G["phi"])
    from ow3d import problems
    x, y = problems.coords(nx, ny, L, L)
    X, Y = np.meshgrid(x, y)
    z = np.sin(np.pi * np.arange(nz) / (2 * (nz - 1)))
    z = np.concatenate([[-z[1]], z])
    m0 = T.store_kinematics(str(tmp_path), 1, 0, G["win"], float(G["dt"]),
                            grid=dict(x=X, y=Y, h=G["h"], hx=G["hx"], hy=G["hy"], z=z))
    assert open(m0, "rb").read() == G["file0"].tobytes()
    m1 = T.store_kinematics(str(tmp_path), 1, 1, G["win"], float(G["dt"]))
    mine, ref = _records(open(m1, "rb").read()), _records(G["file1"].tobytes())
    assert len(mine) == len(ref) == 16
    for q, (a_, b_) in enumerate(zip(mine, ref)):
        av, bv = np.frombuffer(a_, dtype=np.float64), np.frombuffer(b_, dtype=np.float64)
        assert av.shape == bv.shape
        assert np.max(np.abs(av - bv)) <= 1e-12 * max(np.max(np.abs(bv)), 1e-300), q
    T.close()

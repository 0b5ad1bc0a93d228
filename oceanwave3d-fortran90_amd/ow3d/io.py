"""State formats of the reference, so that a tank can be hot-started from a reference run and a
reference run from this one:

* ``OceanWave3D.init`` / ``OceanWave3D.end`` (hot start, IC = -1): a header line, a line
  ``Lx Ly Nx Ny time``, then ``eta Phi`` per interior node, x fastest
  (read: IO/ReadInputFileParameters.f90:112-127 and initialization/SetupInitialConditions.f90:19-27;
  written: main/OceanWave3DTakeATimeStep.f90:146-163).
* Bottom file (Lz <= 0): a header line, a flag line, then per interior node either ``h`` (flag 1) or
  ``h h_x h_xx h_y h_yy`` (flag 0); ghost nodes copy the adjacent boundary values
  (initialization/SetupInitialConditions.f90:51-111).
* ``EP_#####.bin``: Fortran unformatted sequential records ``(nx, ny)``, ``(x, y)``, ``(E, P)`` with
  4-byte record markers, arrays over the ghost-extended plane, x fastest (IO/StoreData.f90:13-29).

Plane arrays here are numpy arrays of shape (nyg, nxg) with the ghost layer, as everywhere in the
package; files hold interior nodes only where the reference's do.
"""
import os

import numpy as np


def _numbers(line, count):
    """The first `count` numbers of a line read list-directed (trailing text such as the reference's
    '<- h,h_x,...' annotations is ignored, as a Fortran list-directed READ ignores it)."""
    vals = []
    for tok in line.replace(",", " ").split():
        try:
            vals.append(float(tok.replace("d", "e").replace("D", "e")))
        except ValueError:
            break
        if len(vals) == count:
            break
    if len(vals) < count:
        raise ValueError(f"expected {count} numbers in line: {line.strip()[:80]}")
    return vals


def _rows(f, n, count):
    out = np.empty((n, count))
    r = 0
    for line in f:
        if not line.strip():
            continue
        out[r] = _numbers(line, count)
        r += 1
        if r == n:
            break
    if r != n:
        raise ValueError(f"expected {n} data lines, found {r}")
    return out


def _interior(nx, ny):
    gy = 1 if ny > 1 else 0
    return slice(gy, gy + ny), slice(1, 1 + nx)


def write_hotstart(path, E, P, lx, ly, time, header="Initial conditions outputted from a previous simulation."):
    """Writes OceanWave3D.end (the reference's restart file) from ghost-extended E, P."""
    nyg, nxg = E.shape
    nx = nxg - 2
    ny = nyg - 2 if nyg > 1 else 1
    sj, si = _interior(nx, ny)
    with open(path, "w") as f:
        f.write(" " + header + "\n")
        f.write(f" {float(lx)!r} {float(ly)!r} {nx} {ny} {float(time)!r}\n")
        for e, p in zip(E[sj, si].ravel(), P[sj, si].ravel()):
            f.write(f" {float(e)!r} {float(p)!r}\n")


def read_hotstart(path, nx, ny):
    """Reads OceanWave3D.init; returns (E, P, info) with E, P ghost-extended (ghost values zero: the
    solver sets them from the boundary conditions) and info = (Lx, Ly, time)."""
    with open(path) as f:
        f.readline()
        lx, ly, nxi, nyi, t0 = _numbers(f.readline(), 5)
        if int(nxi) != nx or int(nyi) != ny:
            raise ValueError(f"initial-condition grid {int(nxi)}x{int(nyi)} does not match the tank {nx}x{ny}")
        data = _rows(f, nx * ny, 2)
    nxg, nyg = nx + 2, (ny + 2 if ny > 1 else 1)
    E = np.zeros((nyg, nxg)); P = np.zeros((nyg, nxg))
    sj, si = _interior(nx, ny)
    E[sj, si] = data[:, 0].reshape(ny, nx)
    P[sj, si] = data[:, 1].reshape(ny, nx)
    return E, P, (lx, ly, t0)


def write_bottom(path, bath, nx, ny, derivatives=True, header="bottom"):
    """Writes a bottom file with h (flag 1) or h, h_x, h_xx, h_y, h_yy (flag 0)."""
    sj, si = _interior(nx, ny)
    with open(path, "w") as f:
        f.write(header + "\n")
        f.write(" 0\n" if derivatives else " 1\n")
        cols = ("h", "hx", "hxx", "hy", "hyy") if derivatives else ("h",)
        arr = np.stack([np.asarray(bath[c])[sj, si].ravel() for c in cols], axis=1)
        for row in arr:
            f.write(" " + " ".join(repr(float(v)) for v in row) + "\n")


def read_bottom(path, nx, ny):
    """Reads a bottom file; returns the ghost-extended dict h, hx, hy, hxx, hyy.  Flag 1 files (h only)
    need numerical derivatives, which the reference computes with its own stencils: not supported."""
    with open(path) as f:
        f.readline()
        flag = int(_numbers(f.readline(), 1)[0])
        if flag != 0:
            raise ValueError("bottom files with h only (flag 1) need numerically computed derivatives: not supported")
        data = _rows(f, nx * ny, 5)
    nxg, nyg = nx + 2, (ny + 2 if ny > 1 else 1)
    out = {}
    sj, si = _interior(nx, ny)
    for c, name in enumerate(("h", "hx", "hxx", "hy", "hyy")):
        a = np.zeros((nyg, nxg))
        a[sj, si] = data[:, c].reshape(ny, nx)
        # ghost nodes copy the adjacent boundary node (SetupInitialConditions.f90:92-108)
        a[:, 0] = a[:, 1]; a[:, -1] = a[:, -2]
        if nyg > 1:
            a[0, :] = a[1, :]; a[-1, :] = a[-2, :]
        out[name] = a
    return out


def _record(f, *arrays):
    payload = b"".join(np.ascontiguousarray(a).tobytes() for a in arrays)
    n = np.array([len(payload)], dtype=np.int32).tobytes()
    f.write(n + payload + n)


def write_ep(path, x, y, E, P):
    """EP_#####.bin as StoreData writes it: (nx, ny), (x, y), (E, P) over the ghost-extended plane;
    x and y are the node coordinates as plane arrays (Fortran order = x fastest)."""
    nyg, nxg = E.shape
    with open(path, "wb") as f:
        _record(f, np.array([nxg, nyg], dtype=np.int32))
        _record(f, np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))
        _record(f, np.asarray(E, dtype=np.float64), np.asarray(P, dtype=np.float64))


def read_ep(path):
    """Reads EP_#####.bin; returns (x, y, E, P) as (nyg, nxg) arrays."""
    def rec(f):
        n = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
        data = f.read(n)
        if int(np.frombuffer(f.read(4), dtype=np.int32)[0]) != n:
            raise ValueError("corrupt record marker")
        return data
    with open(path, "rb") as f:
        nxg, nyg = np.frombuffer(rec(f), dtype=np.int32)
        xy = np.frombuffer(rec(f), dtype=np.float64)
        ep = np.frombuffer(rec(f), dtype=np.float64)
    m = nxg * nyg
    sh = (nyg, nxg)
    return xy[:m].reshape(sh), xy[m:].reshape(sh), ep[:m].reshape(sh), ep[m:].reshape(sh)


def kinematics_path(directory, io, it):
    """Kinematics_<io:2>_<it:5>.bin, the name StoreKinematicData.f90:91-96 gives formattype-21 files."""
    return os.path.join(directory, "Kinematics_%02d_%05d.bin" % (io, it))


def write_kinematics(directory, io, it, win, dt, grid=None, fields=None):
    """Writes one formattype-21 kinematics file as StoreKinematicData.f90:90-478 does (unformatted sequential
    records) and returns its path.  win = (xbeg, xend, xstride, ybeg, yend, ystride, tbeg, tend, tstride) with
    the reference's 1-based interior node numbers.
    it == 0: the grid file -- the window and time stride, dt and the column depth; x, y, h, h_x, h_y of every
    window column; the sigma levels.  grid = dict(x, y, h, hx, hy (ghost-extended planes), z (sigma levels)).
    it > 0: eta, eta_x, eta_y (zero when the grid has one column / row) of the window, then phi, u, v, w and the
    9 gradient fields (ow3d_kinematics_gradients order) of its columns.  fields = dict(E, Ex, Ey (planes),
    phi, U, V, W (3D, W the vertical velocity), G (9 x 3D))."""
    xb, xe, xs, yb, ye, ys, tb, te, ts = (int(v) for v in win)
    ref = grid if it == 0 else fields
    nyg, nxg = np.asarray(ref["h" if it == 0 else "E"]).shape
    gx, gy = 1, (1 if nyg > 1 else 0)
    si = slice(xb + gx - 1, xe + gx, xs)
    sj = slice(yb + gy - 1, ye + gy, ys)
    path = kinematics_path(directory, io, it)
    with open(path, "wb") as f:
        if it == 0:
            nz = len(grid["z"])
            _record(f, np.array([xb, xe, xs, yb, ye, ys, tb, te, ts], dtype=np.int32), np.array([dt], dtype=np.float64),
                    np.array([nz], dtype=np.int32))
            cols = np.stack([np.asarray(grid[k], dtype=np.float64)[sj, si] for k in ("x", "y", "h", "hx", "hy")], axis=-1)
            _record(f, cols)
            _record(f, np.asarray(grid["z"], dtype=np.float64))
            return path
        E = np.asarray(fields["E"], dtype=np.float64)
        _record(f, E[sj, si])
        _record(f, np.asarray(fields["Ex"], dtype=np.float64)[sj, si] if nxg > 1 else np.zeros_like(E[sj, si]))
        _record(f, np.asarray(fields["Ey"], dtype=np.float64)[sj, si] if nyg > 1 else np.zeros_like(E[sj, si]))
        vols = [fields["phi"], fields["U"], fields["V"], fields["W"]] + list(fields["G"])
        for a in vols:
            _record(f, np.asarray(a, dtype=np.float64)[sj, si, :])
    return path

"""Tile and window arithmetic of two multigrid kernels, restated on the CPU and checked for every level size
the hierarchies produce (odd interior counts 3 .. 513 in x and y, wall ghost layers gx = gy = 1):

* k_mg_residual_restrict<T, NZ, 8, 7, 256> (csrc/ow3d_kernels.hpp): a block owns 8 x 7 coarse columns and
  computes the fine residual on the 17 x 15 fine columns their restriction stencils reach.  Every column a
  residual stencil reads (residual_column: 3-point rows, shifted windows at walls, the level-0 shift next to
  a wall) must lie in the staged x window, every fine column a coarse point restricts from must be one the
  block computed, and the fine column of coarse column ic must be 2 ic - 1 (k_mg_restrict_tab's).
* k_mg_zebra_rb modes 3 / 4 (prolongation inside the sweep): the staged coarse window [cxo, cxo + ccols) x
  [cyo, cyo + crows) must hold every coarse column the interpolation tables name for the window's interior
  columns, and fit the static coarse window (TX + 4) / 2 + 2 by (TY + 4) / 2 + 2.
"""
import pytest

SIZES = [3, 5, 9, 17, 33, 65, 129, 257, 513]


def _stencil_cols(i, n):
    """Columns residual_column reads along one axis for column i of n (walls at 0 and n - 1)."""
    if i == 0 or i == n - 1:
        ib = 1 if i == 0 else n - 2
        ws = min(max(ib - 1, 0), n - 3)
        return set(range(ws, ws + 3)) | {i}
    ws = min(max(i - 1, 0), n - 3)
    cols = set(range(ws, ws + 3))
    sh = 1 if ws == 0 else (-1 if ws + 2 == n - 1 else 0)
    if sh:
        cols |= set(range(ws + sh, ws + sh + 3))
    return cols


def _interp_table(nf_in):
    """Host interpolation table (mg_build): per fine index, (left, right) coarse index, None on ghosts."""
    out = []
    for i in range(nf_in + 2):
        if i < 1 or i > nf_in:
            out.append(None)
            continue
        ie = i - 1
        out.append((1 + ie // 2, 1 + ie // 2) if ie % 2 == 0 else (1 + (ie - 1) // 2, 2 + (ie - 1) // 2))
    return out


@pytest.mark.parametrize("nxc", [2, 3, 5, 9, 17, 33, 65, 129, 257])
def test_residual_restrict_tiles_cover_their_stencils(nxc):
    CX, CY, NT = 8, 7, 256
    FW, FH = 2 * CX + 1, 2 * CY + 1
    WX, WY = FW + 4, FH + 4
    for nyc in [2, 5, 17, 65, 257]:
        nxgc, nygc = nxc + 2, nyc + 2
        nxg, nyg = 2 * nxc + 1, 2 * nyc + 1          # fine interior 2 n - 1 plus two ghosts
        for bx in range((nxgc + CX - 1) // CX):
            for by in range((nygc + CY - 1) // CY):
                ic0, jc0 = bx * CX, by * CY
                fi0, fj0 = 2 * ic0 - 2, 2 * jc0 - 2
                lox, loy = max(min(fi0 - 2, nxg - 4), 0), max(min(fj0 - 2, nyg - 4), 0)
                wc, wr = min(WX, nxg - lox), min(WY, nyg - loy)
                computed = set()
                for t in range(NT):
                    tx, ty = t % FW, t // FW
                    i, j = fi0 + tx, fj0 + ty
                    if t < FW * FH and 0 <= i < nxg and 0 <= j < nyg:
                        computed.add((tx, ty))
                        assert all(lox <= c < lox + wc for c in _stencil_cols(i, nxg))
                        assert all(loy <= c < loy + wr for c in _stencil_cols(j, nyg))
                for rr in range(CY):
                    for cc in range(CX):
                        ic, jc = ic0 + cc, jc0 + rr
                        if ic >= nxgc or jc >= nygc or not (1 <= ic <= nxc and 1 <= jc <= nyc):
                            continue
                        ti, tj = 2 * cc + 1, 2 * rr + 1
                        assert fi0 + ti == 2 * ic - 1 and fj0 + tj == 2 * jc - 1
                        for a in (ti - 1, ti, ti + 1):
                            for b in (tj - 1, tj, tj + 1):
                                assert (a, b) in computed


@pytest.mark.parametrize("TX,TY", [(16, 8), (32, 4), (16, 4), (8, 4)])
def test_prolongating_sweep_coarse_window_holds_every_interpolation_source(TX, TY):
    WX, WY = TX + 4, TY + 4
    CWX, CWY = WX // 2 + 2, WY // 2 + 2
    for nxf in SIZES:
        for nyf in SIZES:
            nxg, nyg = nxf + 2, nyf + 2
            nxc, nyc = (nxf + 1) // 2, (nyf + 1) // 2
            tx, ty = _interp_table(nxf), _interp_table(nyf)
            offx = 1 if nxf % TX == 1 else 0
            offy = 1 if nyf % TY == 1 else 0
            for bx in range((nxf + offx + TX - 1) // TX):
                for by in range((nyf + offy + TY - 1) // TY):
                    i0 = 1 + max(0, bx * TX - offx)
                    ie = 1 + min(nxf, (bx + 1) * TX - offx)
                    j0 = 1 + max(0, by * TY - offy)
                    je = 1 + min(nyf, (by + 1) * TY - offy)
                    lox, loy = max(i0 - 2, 0), max(j0 - 2, 0)
                    wcols, wrows = min(ie + 2, nxg) - lox, min(je + 2, nyg) - loy
                    assert wcols <= WX and wrows <= WY
                    cxo, cyo = ((max(lox, 1) - 1) >> 1) + 1, ((max(loy, 1) - 1) >> 1) + 1
                    ccols = (min(lox + wcols - 1, nxg - 2) >> 1) + 2 - cxo
                    crows = (min(loy + wrows - 1, nyg - 2) >> 1) + 2 - cyo
                    assert 1 <= ccols <= CWX and 1 <= crows <= CWY
                    assert cxo + ccols <= nxc + 2 and cyo + crows <= nyc + 2
                    for c in range(wcols):
                        e = tx[lox + c]
                        if e is not None:
                            assert cxo <= e[0] and e[1] < cxo + ccols
                    for r in range(wrows):
                        e = ty[loy + r]
                        if e is not None:
                            assert cyo <= e[0] and e[1] < cyo + crows

"""Golden vectors written by the reference's own routines (tests/golden/make_golden.py, run against
oracle/_ref compiled from the reference sources).  The CPU tests pin the oracle to them bit for bit;
the GPU tests hold the HIP path to them at the FP64 tolerance of the product (<= 1e-12 relative L-inf
for the matvec, plane derivatives, W and the free-surface rhs; <= 1e-14 for the ghost updates).
"""
import glob
import os

import numpy as np
import pytest

from conftest import oracle_params

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(os.path.basename(p)[4:-4] for p in glob.glob(os.path.join(HERE, "ref_*.npz")) if "transfer" not in p)


def load(name):
    with np.load(os.path.join(HERE, f"ref_{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def params(G, **kw):
    a = int(G["alpha"])
    return oracle_params(int(G["nx"]), int(G["ny"]), int(G["nz"]), alpha=a, beta=a, gamma=a, gridz=int(G["gridz"]),
                         lx=float(G["L"]), ly=float(G["L"]), precond=0, **kw)


def test_fixtures_present():
    assert len(CASES) >= 4 and os.path.exists(os.path.join(HERE, "ref_transfer17x17x9.npz"))


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden_vectors(oracle_mod, name):
    G = load(name)
    T = oracle_mod.OracleTank(params(G), G["h"], G["hx"], G["hy"], G["hxx"], G["hyy"])
    x, y, z = T.grid()
    assert np.array_equal(x, G["x"]) and np.array_equal(y, G["y"]) and np.array_equal(z, G["z"])
    sx, sy, sz = T.stencils()
    assert np.array_equal(sx, G["sx"]) and np.array_equal(sz, G["sz"])
    if int(G["ny"]) > 1:
        assert np.array_equal(sy, G["sy"])
    E, P, d = T.surface(G["E_in"], G["P_in"])
    assert np.array_equal(E, G["E"]) and np.array_equal(P, G["P"])
    for k in ("Ex", "Exx", "Px") + (("Ey", "Eyy", "Py") if int(G["ny"]) > 1 else ()):
        assert np.array_equal(d[k], G[k]), k
    assert np.array_equal(T.dsigma(), G["dsigma"])
    assert np.array_equal(T.matvec(G["phi"]), G["Aphi"])
    W = T.vertical_velocity(G["phi"])
    assert np.array_equal(W, G["W"])
    rE, rP = T.rhs(E, P, G["W"])
    assert np.array_equal(rE, G["rhsE"]) and np.array_equal(rP, G["rhsP"])
    T.close()


def test_oracle_transfers_match_reference_golden_vectors(oracle_mod):
    G = load("transfer17x17x9")
    T = oracle_mod.OracleTank(oracle_params(17, 17, 9, lx=10.0, ly=10.0), G["h"], G["hx"], G["hy"], G["hxx"], G["hyy"])
    assert T.nlev == int(G["nlev"])
    for fine in range(1, T.nlev):
        assert np.array_equal(T.restrict(fine, G[f"rf{fine}"]), G[f"rc{fine}"])
        assert np.array_equal(T.prolong(fine, G[f"ec{fine}"]), G[f"ef{fine}"])
    T.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_reference_golden_vectors(name):
    import ow3d
    G = load(name)
    p = params(G)
    cfg = {k: v for k, v in p.items() if k not in ("mg_mode", "cd_quirk")}
    T = ow3d.Tank(cfg, G["h"], G["hx"], G["hy"], G["hxx"], G["hyy"])
    E, P, d = T.set_surface(G["E_in"], G["P_in"])
    assert rel(E, G["E"]) <= 1e-14 and rel(P, G["P"]) <= 1e-14
    for k in ("Ex", "Exx", "Px") + (("Ey", "Eyy", "Py") if int(G["ny"]) > 1 else ()):
        assert rel(d[k], G[k]) <= 1e-12, k
    assert rel(T.laplace_apply(G["phi"]), G["Aphi"]) <= 1e-12
    W = T.vertical_velocity(G["phi"])
    assert rel(W, G["W"]) <= 1e-12
    rE, rP = T.rhs_free_surface(E, G["W"])
    assert rel(rE, G["rhsE"]) <= 1e-12 and rel(rP, G["rhsP"]) <= 1e-12
    T.close()

"""Host logic of the many-tank and request-loop paths (no GPU): the workload draws, the Poisson request
stream, the device-memory estimate used for admission, and the even-Nz coarsening rule of the
multigrid hierarchy (on the oracle, which follows the same rule)."""
import numpy as np
import pytest


def test_batch_draws_respect_the_workload_ranges():
    from ow3d import problems
    seen = set()
    for s in range(300):
        d = problems.batch_draw(s)
        assert d["nx"] in problems.BATCH_SIZES and d["ny"] in problems.BATCH_SIZES and d["nz"] in problems.BATCH_NZ
        assert d["order"] in problems.BATCH_ORDERS and d["order"] + 1 <= d["nz"] + 1   # sigma stencil fits the column
        assert 0.5 <= d["h0"] <= 2.0 and 0.5 <= d["kph"] <= 3.0 and 0.005 <= d["hs"] / d["h0"] <= 0.03
        cfg = problems.batch_cfg(d)
        lam = 2 * np.pi * d["h0"] / d["kph"]
        assert abs(cfg["lx"] / (d["nx"] - 1) - lam / 16) < 1e-12 * lam   # 16 points per peak wavelength
        assert cfg["gmres_m"] == (31 if d["order"] == 8 else 23)   # 8th order converges in GMRES(31)
        seen.add((d["nz"], d["order"]))
    assert (6, 8) not in seen and (12, 8) in seen and (6, 4) in seen
    # a draw is a function of its seed
    assert problems.batch_draw(7) == problems.batch_draw(7)


def test_poisson_request_stream():
    from ow3d import serve
    reqs = serve.poisson_requests(2000, rate=20.0, seed=3)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.1 / 20.0
    steps = np.array([r.nsteps for r in reqs])
    assert steps.min() >= 50 and steps.max() <= 500
    bottoms = {r.draw["bottom"] for r in reqs}
    assert bottoms == {"flat", "shoal"}
    assert all(r.bytes > 0 and r.dof == r.cfg["nx"] * r.cfg["ny"] * (r.cfg["nz"] + 1) for r in reqs)


def test_device_memory_estimate():
    import ow3d
    base = dict(nx=129, ny=129, nz=9, lx=10.0, ly=10.0)
    e64 = ow3d.estimate_bytes(dict(base, mg_precision=0))
    e32 = ow3d.estimate_bytes(dict(base, mg_precision=1))
    e15 = ow3d.estimate_bytes(dict(base, mg_precision=1, gmres_m=15))
    ng = 131 * 131 * 10
    assert e32 < e64 and e15 < e32
    assert e32 - e15 == 8 * 8 * ng              # eight fewer Krylov basis vectors
    # flexible GMRES keeps m preconditioned vectors more
    assert ow3d.estimate_bytes(dict(base, mg_precision=1, gmres_side=1)) - e32 == 23 * 8 * ng
    # fine-grid vectors + basis (m = 23) + step snapshot dominate: about (8 + 24 + 1) vectors
    assert 33 * 8 * ng < e64 < 45 * 8 * ng
    with pytest.raises(ow3d.OW3DError):
        ow3d.estimate_bytes(dict(base, nz=2))


def test_even_columns_are_kept_while_the_plane_is_coarsened(oracle_mod):
    from conftest import oracle_params
    from ow3d import problems
    for (nx, ny, nz, want) in [(17, 17, 12, [(17, 17, 12), (9, 9, 12), (5, 5, 12), (3, 3, 12)]),
                               (33, 17, 6, [(33, 17, 6), (17, 9, 6), (9, 5, 6), (5, 5, 6), (3, 3, 6)]),
                               (17, 17, 9, [(17, 17, 9), (9, 9, 9), (5, 5, 5), (3, 3, 3)])]:
        bath = problems.flat_bottom(nx, ny, 5.0, 5.0)
        p = oracle_params(nx, ny, nz, alpha=1, beta=1, gamma=1, lx=5.0, ly=5.0, mg_maxgrids=0)
        T = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
        got = [(a - 2, b - 2, c - 1) for (a, b, c) in (T.level_dims(k) for k in range(T.nlev - 1, -1, -1))]
        assert got[:len(want)] == want, (nx, ny, nz, got)


def test_autotuner_candidates_and_cache(tmp_path):
    import ow3d
    from ow3d import autotune
    for cfg in (dict(nx=513, ny=513, nz=9), dict(nx=65, ny=129, nz=12), dict(nx=129, ny=1, nz=9)):
        cands = autotune.level_candidates(cfg)
        assert cands == sorted(cands) and 1 <= len(cands) <= 3
        for L in cands:
            n, nc = ow3d.mg_plan(dict(cfg, mg_maxgrids=L))
            assert n == L and nc <= autotune.MAX_COARSE
        # one level fewer than the first candidate would exceed the dense-solve limit
        if cands[0] > 2:
            assert ow3d.mg_plan(dict(cfg, mg_maxgrids=cands[0] - 1))[1] > autotune.MAX_COARSE
    assert ow3d.mg_plan(dict(nx=513, ny=513, nz=9, mg_maxgrids=7)) == (7, 11 * 11 * 10)   # 9^2 x 9 coarsest
    path = str(tmp_path / "t.json")
    table = {autotune.class_key(dict(nx=65, ny=65, nz=9, alpha=2)): {"mg_maxgrids": 4, "matvec_variant": 1, "sweep_threads": 0}}
    autotune.save(table, path)
    back = autotune.load(path)
    assert autotune.tuned(dict(nx=65, ny=65, nz=9, alpha=2), back)["matvec_variant"] == 1
    assert autotune.tuned(dict(nx=65, ny=65, nz=9, alpha=3), back) is None


def test_requests_larger_than_the_budget_are_rejected_on_arrival():
    """A request whose tank cannot fit even the whole HBM budget is rejected when it arrives rather than
    blocking the queue: with every request too large the loop returns at once, having built nothing."""
    from ow3d import serve
    reqs = serve.poisson_requests(3, rate=1000.0, seed=2, mg_precision=1, steps=(2, 3), sizes=(65,))
    out = serve.Server(hbm_budget=min(r.bytes for r in reqs) / 2, max_active=4).run(reqs, time_limit=None)
    assert sorted(out["rejected"]) == [r.rid for r in reqs]
    assert out["admitted"] == 0 and out["completed"] == 0 and out["steps"] == 0

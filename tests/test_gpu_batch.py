"""Many tanks at once on the GPU: device-driven solves, captured RK4 steps, batched and asynchronous
stepping.  Tanks share nothing on the device (each has its own buffers, stream, reductions and captured
graphs), so a tank's result must be bit-identical whether it runs alone, in a batch of uneven tanks, or
through the asynchronous entry points; and a replayed step graph must give bit-identical results to the
same step launched eagerly."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = (17, 33, 65)


def _ow3d():
    import ow3d
    return ow3d


def make(spec, mg_precision=1, eager=False):
    ow3d = _ow3d()
    cfg, bath, E, P, dt, zones = spec
    c = dict(cfg)
    c.update(mg_precision=mg_precision, reltol=1e-8, gmres_m=23)
    if eager:
        os.environ["OW3D_NO_GRAPH"] = "1"
    try:
        T = ow3d.Tank(c, **bath)
    finally:
        if eager:
            del os.environ["OW3D_NO_GRAPH"]
    for z in zones:
        T.add_relax_zone(*z)
    T.set_state(E, P)
    return T, dt


def specs(seeds, sizes=SMALL):
    from ow3d import problems
    return [problems.batch_tank(s, sizes=sizes) for s in seeds]


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_is_bit_identical_to_each_tank_alone(prec):
    ow3d = _ow3d()
    sp = specs([11, 12, 13, 14, 15])
    print("batch tanks:", [(s[0]["nx"], s[0]["ny"], s[0]["nz"], 2 * s[0]["alpha"]) for s in sp], flush=True)
    nsteps = 4
    alone, status = [], 0
    for s in sp:
        T, dt = make(s, prec)
        # order 8 with the 2nd-order preconditioner can need more than GMRES(23) allows at rtol 1e-8
        # (the oracle stops at 23 matvecs on these tanks too): status 1 is reported, not an error
        st = T.rk4_steps(nsteps, 0.0, dt)
        assert st in (0, 1)
        status |= st
        alone.append(T.get_state())
        T.close()
    tanks = [make(s, prec) for s in sp]
    st = ow3d.batch_rk4_steps([t for t, _ in tanks], nsteps, [dt for _, dt in tanks])
    assert st == status
    for (T, _), ref in zip(tanks, alone):
        assert same(T.get_state(), ref)
        T.close()


def test_async_steps_equal_synchronous_steps():
    sp = specs([21, 22])
    A = [make(s) for s in sp]
    B = [make(s) for s in sp]
    sts = [T.rk4_steps(5, 0.0, dt) for T, dt in A]
    for T, dt in B:
        T.rk4_steps_async(5, dt)
    for (T, _), st in zip(B, sts):
        assert T.wait() == st and T.query()
    for (Ta, _), (Tb, _) in zip(A, B):
        assert same(Ta.get_state(), Tb.get_state())


def test_step_graph_replays_equal_eager_steps():
    """Each RK4 step from the second on replays a captured graph; the replays, with host work and
    synchronisation between them, must match eager launches of the same steps bit for bit."""
    from ow3d import problems
    spec = problems.batch_tank(31, sizes=(33,), nzs=(9,), orders=(6,), bottom="shoal")
    Tg, dt = make(spec)
    Te, _ = make(spec, eager=True)
    for step in range(6):
        Tg.rk4_steps(1, 0.0, dt)
        Te.rk4_steps(1, 0.0, dt)
        assert same(Tg.get_state(), Te.get_state()), step
    assert Tg.stats() == Te.stats()


def test_even_nz_tanks_build_their_multigrid_without_z_coarsening():
    from ow3d import problems
    for nz in (6, 12):
        spec = problems.batch_tank(41, sizes=(65,), nzs=(nz,), orders=(4,), bottom="flat")
        T, dt = make(spec)
        assert T.mg_levels >= 3
        assert T.rk4_steps(3, 0.0, dt) == 0
        tot, solves, mn, mx = T.stats()
        assert solves == 12 and mx <= 23


@pytest.mark.parametrize("prec", [0, 1])
def test_estimated_device_bytes_match_the_allocation(prec):
    ow3d = _ow3d()
    for s in specs([51, 52, 53], sizes=(33, 65, 129)):
        cfg = dict(s[0]); cfg.update(mg_precision=prec, gmres_m=23)
        T, _ = make(s, prec)
        est, got = ow3d.estimate_bytes(cfg), T.device_bytes()
        assert 0.98 * got <= est <= 1.02 * got, (est, got)
        T.close()


def test_footprint_of_a_handle_created_while_another_is_destroyed():
    """A handle's device_bytes counts its own buffers only: destroying another handle on another thread
    while it is being created (creation releases the capture lock during its host-side coarse inverse)
    must not change it."""
    import threading
    from ow3d import problems
    spec = problems.batch_tank(81, sizes=(129,), nzs=(9,), orders=(6,), bottom="shoal")
    ref, _ = make(spec)
    want = ref.device_bytes()
    ref.close()
    others = [make(s)[0] for s in specs([82, 83, 84], sizes=(65,))]
    out = {}

    def build():
        out["T"] = make(spec)[0]

    th = threading.Thread(target=build)
    th.start()
    for T in others:
        T.close()
    th.join()
    assert out["T"].device_bytes() == want
    out["T"].close()


def test_even_nz_multigrid_matches_the_oracle(oracle_mod):
    """Nz = 12: the multigrid keeps the even column and halves x and y down to 3 points (17, 9, 5, 3);
    the oracle follows the same rule and must agree with the GPU over RK4 steps."""
    from conftest import oracle_params
    from ow3d import problems
    spec = problems.batch_tank(11, sizes=(17,), nzs=(12,), orders=(6,), bottom="shoal")
    cfg, bath, E, P, dt, _ = spec
    p = oracle_params(cfg["nx"], cfg["ny"], cfg["nz"], alpha=3, beta=3, gamma=3, lx=cfg["lx"], ly=cfg["ly"], mg_maxgrids=0)
    T_o = oracle_mod.OracleTank(p, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])
    c = {k: v for k, v in p.items() if k not in ("mg_mode", "cd_quirk")}
    T_g = _ow3d().Tank(c, **bath)
    assert T_g.mg_levels == T_o.nlev == 4
    T_o.set_state(E, P)
    T_g.set_state(E, P)
    for s in range(4):
        T_o.rk4_step(s * dt, dt)
    T_g.rk4_steps(4, 0.0, dt)
    Eo, Po, _, _ = T_o.get_state()
    Eg, Pg, _, _ = T_g.get_state()
    assert np.max(np.abs(Eg - Eo)) <= 1e-9 * np.max(np.abs(Eo))
    assert np.max(np.abs(Pg - Po)) <= 1e-9 * np.max(np.abs(Po))
    assert T_g.stats()[0] == T_o.stats()[0]


def test_speculative_steps_replay_to_the_whole_budget_result():
    """A tank whose solves are enqueued with too few iteration slots (margin -3: three fewer than the
    most any solve of that stage used in the last steps) is rolled back and replayed with the whole
    budget; the result must equal stepping with the whole budget every time, bit for bit."""
    from ow3d import problems
    spec = problems.batch_tank(61, sizes=(33,), nzs=(9,), orders=(6,), bottom="shoal")
    os.environ["OW3D_SPECULATE"] = "0"
    try:
        Tf, dt = make(spec)
    finally:
        del os.environ["OW3D_SPECULATE"]
    os.environ["OW3D_SLOT_MARGIN"] = "-3"
    try:
        Ts, _ = make(spec)
    finally:
        del os.environ["OW3D_SLOT_MARGIN"]
    for step in range(8):
        Tf.rk4_steps(1, 0.0, dt)
        Ts.rk4_steps(1, 0.0, dt)
        assert same(Tf.get_state(), Ts.get_state()), step
    assert Tf.stats() == Ts.stats()
    assert Tf.replays() == 0
    assert Ts.replays() >= 6, Ts.replays()


def test_request_loop_results_equal_each_request_alone():
    """Requests arriving while others run are admitted as slots free up (2 slots for 6 requests), their
    steps overlap on the device, and every request's final state equals its run alone, bit for bit."""
    from ow3d import serve
    reqs = serve.poisson_requests(6, rate=200.0, seed=5, mg_precision=1, steps=(3, 9), sizes=(17, 33, 65))
    srv = serve.Server(hbm_budget=50e9, max_active=2, keep_results=True)
    out = srv.run(reqs)
    assert out["completed"] == 6 and out["admitted"] == 6
    assert out["steps"] == sum(r.nsteps for r in reqs)
    assert out["queue_delay_s"]["p99"] > 0.0          # some requests waited for a slot
    for r in reqs:
        assert same(out["results"][r.rid], serve.run_alone(r)), r.rid


def test_request_loop_reports_unconverged_solves():
    """Solves limited to GMRES(6) at rtol 1e-14 cannot reach their tolerance: the loop must count every
    such step per request, as ow3d_steps_status reports it."""
    from ow3d import serve
    reqs = serve.poisson_requests(3, rate=1000.0, seed=9, mg_precision=1, steps=(2, 3), sizes=(33,))
    for r in reqs:
        r.cfg.update(gmres_m=6, reltol=1e-14)
    out = serve.Server(hbm_budget=50e9, max_active=3).run(reqs)
    assert out["completed"] == 3
    assert out["completed_with_unconverged_solves"] == 3
    assert out["unconverged_steps_of_completed"] == out["steps"]


def test_request_loop_respects_the_memory_budget():
    from ow3d import serve
    reqs = serve.poisson_requests(4, rate=1000.0, seed=6, mg_precision=1, steps=(2, 3), sizes=(33,))
    budget = max(r.bytes for r in reqs) * 1.5         # room for one tank at a time
    out = serve.Server(hbm_budget=budget, max_active=8).run(reqs)
    assert out["completed"] == 4
    assert out["queue_delay_s"]["p99"] > 0.0


def test_launch_choices_do_not_change_results():
    """Zebra-sweep block size only changes the thread mapping (bit-identical); the other matvec kernels
    sum the same terms in another order (within 1e-11)."""
    from ow3d import problems
    spec = problems.batch_tank(71, sizes=(65,), nzs=(9,), orders=(6,), bottom="shoal")
    ref, dt = make(spec)
    ref.rk4_steps(4, 0.0, dt)
    Eref = ref.get_state()
    for mv, thr, exact in ((0, 256, True), (0, 1024, True), (1, 0, False), (2, 0, False), (3, 0, False)):
        T, _ = make(spec)
        T.tune(mv, thr)
        T.rk4_steps(4, 0.0, dt)
        got = T.get_state()
        if exact:
            assert same(got, Eref), (mv, thr)
        else:
            for a, b in zip(got, Eref):
                assert np.max(np.abs(a - b)) <= 1e-11 * max(np.max(np.abs(b)), 1e-300), (mv, thr)
        T.close()
    with pytest.raises(Exception):
        ref.tune(4, 0)
